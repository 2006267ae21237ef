"""Build libfmmx.so (and libfmmx_ref.so) in place with hipcc for gfx950.

Sources under old_fmm_bem_amd/csrc/ are compiled to objects in _build/ (in
parallel) and linked into old_fmm_bem_amd/libfmmx.so.  The shared library is
what the Python host layer, the tests and the bench load; it is rebuilt from
source on every call.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfmmx.so")
REFLIB = os.path.join(HERE, "libfmmx_ref.so")

ARCH = os.environ.get("FMMX_ARCH", "gfx950")
# packed FP32 instructions are disabled for all device code
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-I" + os.path.join(ROOT, "include"),
          "-Wall", "-Wno-unused-function", "-Wno-unused-result"]

CORE = ["precompute.cpp", "yukawa.cpp", "helmholtz.cpp", "split_host.cpp", "tree.hip", "kernels.hip", "m2l_split.hip",
        "api.cpp", "adjoint.cpp"]
SHIM = ["refshim.cpp"]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _stale(src: str, obj: str) -> bool:
    """The object is older than its source or any header it may include."""
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    deps = [os.path.join(CSRC, src)] + [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith(".h")]
    deps += [os.path.join(ROOT, "include", h) for h in os.listdir(os.path.join(ROOT, "include"))]
    return any(os.path.getmtime(d) > t for d in deps)


def _compile(src: str, incremental: bool = False) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    if incremental and not _stale(src, obj):
        return obj
    cmd = [hipcc(), *COMMON, *DEVICE_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd.insert(1, "-xhip")
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile {src} failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False, incremental: bool = False) -> str:
    """Compile every source (incremental: only objects older than their
    source or a header) and link the libraries."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = CORE + [s for s in SHIM if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = dict(zip(srcs, ex.map(lambda s_: _compile(s_, incremental), srcs)))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *[objs[s] for s in CORE],
           "-Wl,-soname,libfmmx.so"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if "refshim.cpp" in objs:
        tmp = REFLIB + ".tmp"
        cmd = [hipcc(), "-shared", "-fPIC", "-o", tmp, objs["refshim.cpp"], "-L" + HERE, "-lfmmx",
               "-Wl,-rpath,$ORIGIN", "-Wl,-soname,libfmmx_ref.so"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link refshim failed:\n{r.stdout}\n{r.stderr}")
        os.replace(tmp, REFLIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True, incremental="--incremental" in sys.argv)
    sys.exit(0)
