"""Write the golden vectors under tests/golden/ from the reference itself.

Runs only in a container that holds the reference sources: it builds
oracle/_ref/ref_driver with oracle/Makefile.ref (the reference's serial CPU
FMM compiled from its own sources, compile-time mp=9 as shipped in
include/constants.h:12) and feeds it seeded inputs.  Each case is saved as
ref_<name>.npz holding the inputs and every array the driver recorded:
sort permutations, box keys/ranges, near and far interaction lists, precalc
tables, multipoles after P2M and after the upward pass, local expansions per
level, P2P-only and final potentials/gradients (sorted and caller order), the
direct sum.  Yukawa cases (neq 10, the y* names) hold the shipped fmm()'s
potentials and the direct sum kernel/Helmholtz_G_dir.cxx.  ref_box.npz holds point clouds and
the root cube bio/main.cxx:355-368 computes for each (the driver's box
mode).  (The shipped fmm() itself is not run: with its level-2 neib=3
list it reads past the end of nfj at kernel/m2l.cxx:43, a heap overflow
under AddressSanitizer.)

Usage: python tests/golden/gen_golden.py [--ref /root/reference]
"""
from __future__ import annotations

import argparse
import os
import struct
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
ORACLE = os.path.join(ROOT, "oracle")

PI_F = np.float32(np.pi)

# name: (ni, nj, lmax, neq, same_points, asis, seed)
CASES = {
    "g3": (1200, 1100, 3, 0, False, False, 11),
    "gni3": (900, 1000, 3, 1, False, False, 12),
    "gnj3": (1000, 900, 3, 2, False, False, 13),
    "g2": (400, 450, 2, 0, False, False, 14),
    "g4same": (2500, 2500, 4, 0, True, False, 15),
    "gedge4": (1500, 1400, 4, 0, False, False, 16),
    # Yukawa (neq 10), test/testh.cxx:37-42 style: lambda 0.1, same points
    "y2": (1000, 1000, 2, 10, True, False, 17),
    "y2l1": (1500, 1500, 2, 10, True, False, 18),
}
# The shipped Yukawa path runs only at lmax 2: at lmax 3 Helmholtz_m2l indexes
# lexp1/lexp2 (kernel/Helmholtz_m2l.cxx:111-122) with stacked box numbers
# beyond the nbne rows it allocates, and crashes.
LAMBDA = {"y2": 0.1, "y2l1": 1.0}
# arrays kept in every case; expansion dumps only where listed (size)
FULL = {"g3", "gnj3", "g2"}


def make_inputs(ni, nj, same, seed):
    """test/test.cxx:72-92 style: points uniform in [-pi,pi)^3 (float),
    normals/charges uniform in [0,1) -- drawn with numpy, not rand()."""
    rng = np.random.default_rng(seed)
    xi = (rng.random((ni, 3), dtype=np.float32) * np.float32(2) * PI_F - PI_F).astype(np.float32)
    ni_n = rng.random((ni, 3), dtype=np.float32)
    if same:
        xj = xi.copy()
    else:
        xj = (rng.random((nj, 3), dtype=np.float32) * np.float32(2) * PI_F - PI_F).astype(np.float32)
    nj_n = rng.random((nj, 3), dtype=np.float32)
    q = rng.random(nj, dtype=np.float32)
    return xi, ni_n, xj, nj_n, q


def face_nudged(n, lmax, rng):
    """Points whose coordinates sit within 8 ulp of the box faces of level
    lmax: fmm/boxn.cxx:9 scales by kl/rd computed in float, and the keys of
    such points depend on that rounding."""
    kl = 1 << lmax
    rd = np.float32(2) * PI_F
    faces = np.array([np.float32(-PI_F + np.float32(c) * (rd / np.float32(kl))) for c in range(1, kl)], np.float32)
    x = (rng.random((n, 3), dtype=np.float32) * np.float32(2) * PI_F - PI_F).astype(np.float32)
    for k in range(3):
        sel = rng.random(n) < 0.7
        base = faces[rng.integers(0, kl - 1, sel.sum())]
        ulp = rng.integers(-8, 9, sel.sum())
        v = base.copy()
        for i in range(len(v)):
            for _ in range(abs(int(ulp[i]))):
                v[i] = np.nextafter(v[i], np.float32(np.inf) if ulp[i] > 0 else np.float32(-np.inf))
        x[sel, k] = v
    return x


def write_input(path, ni, nj, lmax, neq, asis, xi, ni_n, xj, nj_n, q, lam=None):
    box = np.array([-PI_F, -PI_F, -PI_F, np.float32(2) * PI_F], np.float32)
    with open(path, "wb") as f:
        f.write(struct.pack("<5i", ni, nj, lmax, neq, int(asis)))
        f.write(box.tobytes())
        for k in range(3):
            f.write(np.ascontiguousarray(xi[:, k]).tobytes())
        for k in range(3):
            f.write(np.ascontiguousarray(ni_n[:, k]).tobytes())
        for k in range(3):
            f.write(np.ascontiguousarray(xj[:, k]).tobytes())
        for k in range(3):
            f.write(np.ascontiguousarray(nj_n[:, k]).tobytes())
        f.write(q.tobytes())
        if neq == 10:
            f.write(np.float32(lam).tobytes())
    return box


def read_dump(path):
    out = {}
    data = open(path, "rb").read()
    p = 0
    while p < len(data):
        name = data[p:p + 32].split(b"\0", 1)[0].decode()
        typ, n = struct.unpack_from("<2i", data, p + 32)
        p += 40
        dt = {0: np.int32, 1: np.float32, 2: np.float64}[typ]
        es = np.dtype(dt).itemsize
        out[name] = np.frombuffer(data[p:p + es * n], dtype=dt).copy()
        p += es * n
    return out


def box_clouds():
    """Point clouds for the root cube of bio/main.cxx:355-368: two-point
    clouds whose x extent is a float where rounding the double product
    extent*(1+1e-5) once to float differs from a float product (found by
    search, plus 0.40973523), and seeded random clouds of several shapes."""
    rng = np.random.default_rng(40)
    clouds = []
    ext = [np.float32(0.40973523)]
    while len(ext) < 16:
        e = np.float32(rng.uniform(0.01, 50.0))
        if np.float32(np.float64(e) * (1 + 1e-5)) != np.float32(e * np.float32(1 + 1e-5)):
            ext.append(e)
    for e in ext:
        lo = np.float32(rng.uniform(-3, 3))
        x = np.array([[lo, 0.1, 0.2], [np.float32(lo + e), 0.15, 0.25]], np.float32)
        if np.float32(x[1, 0] - x[0, 0]) != e:  # keep the float extent exact
            x[1, 0] = np.float32(x[0, 0] + e)
        clouds.append(x)
    for k in range(12):
        n = int(rng.integers(50, 2000))
        if k % 3 == 0:
            x = rng.random((n, 3), dtype=np.float32) * np.float32(rng.uniform(0.1, 10)) - np.float32(rng.uniform(0, 5))
        elif k % 3 == 1:
            v = rng.standard_normal((n, 3))
            x = (v / np.linalg.norm(v, axis=1)[:, None] * rng.uniform(0.5, 3)).astype(np.float32)
        else:
            x = (rng.standard_normal((n, 3)) * rng.uniform(0.5, 2, 3)).astype(np.float32)
        clouds.append(x.astype(np.float32))
    return clouds


def write_box_golden(out_dir):
    clouds = box_clouds()
    tmp = os.path.join(ORACLE, "_ref", "tmp")
    os.makedirs(tmp, exist_ok=True)
    fin, fout = os.path.join(tmp, "box.in"), os.path.join(tmp, "box.out")
    with open(fin, "wb") as f:
        f.write(struct.pack("<i", len(clouds)))
        for x in clouds:
            f.write(struct.pack("<i", len(x)))
            for k in range(3):
                f.write(np.ascontiguousarray(x[:, k], np.float32).tobytes())
    subprocess.run([os.path.join(ORACLE, "_ref", "ref_driver"), "box", fin, fout], check=True)
    d = read_dump(fout)
    keep = {}
    for k, x in enumerate(clouds):
        keep[f"cloud_{k}"] = x
        keep[f"box_{k}"] = d[f"box_{k}"]
    dst = os.path.join(out_dir, "ref_box.npz")
    np.savez_compressed(dst, **keep)
    print(f"box: {len(clouds)} clouds -> {os.path.relpath(dst, ROOT)}")
    os.remove(fin)
    os.remove(fout)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("--only", default=None)
    ap.add_argument("--exe", default=None, help="another build of ref_driver (e.g. with AddressSanitizer)")
    ap.add_argument("--out", default=HERE)
    a = ap.parse_args()
    subprocess.run(["make", "-s", "-C", ORACLE, "-f", "Makefile.ref", f"REF={a.ref}", "-j8"], check=True)
    # The shipped Yukawa P2M writes one element past multipoleh's inr[mpcmp]
    # (kernel/multipoleh.cxx:10-12 -> kernel/in.cxx:33-35: ribesl fills
    # b[0..mp]); with the default stack protector that aborts.  Yukawa cases
    # therefore run a build with AddressSanitizer in recover mode, where the
    # extra element lands in a redzone and nothing else is overwritten (same
    # strict floating-point flags, so the arithmetic is unchanged).
    if not a.only or a.only == "box":
        write_box_golden(a.out)
    if a.only == "box":
        return 0
    asan_flags = "-O1 -fno-fast-math -ffp-contract=off -w -fsanitize=address -fsanitize-recover=address"
    if any(CASES[c][3] == 10 for c in CASES if not a.only or c == a.only):
        subprocess.run(["make", "-s", "-C", ORACLE, "-f", "Makefile.ref", f"REF={a.ref}", "OUT=_ref/asan",
                        f"CXXFLAGS={asan_flags}", "-j8"], check=True)
    asan_exe = os.path.join(ORACLE, "_ref", "asan", "ref_driver")
    asan_env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=0")
    tmp = os.path.join(ORACLE, "_ref", "tmp")
    os.makedirs(tmp, exist_ok=True)
    for name, (ni, nj, lmax, neq, same, asis, seed) in CASES.items():
        if a.only and name != a.only:
            continue
        xi, ni_n, xj, nj_n, q = make_inputs(ni, nj, same, seed)
        if name.startswith("gedge"):
            rng = np.random.default_rng(seed + 1000)
            xi = face_nudged(ni, lmax, rng)
            xj = face_nudged(nj, lmax, rng)
        fin = os.path.join(tmp, f"{name}.in")
        fout = os.path.join(tmp, f"{name}.out")
        box = write_input(fin, ni, nj, lmax, neq, asis, xi, ni_n, xj, nj_n, q, LAMBDA.get(name))
        if neq == 10:
            subprocess.run([a.exe or asan_exe, fin, fout], check=True, stdout=subprocess.DEVNULL,
                           stderr=subprocess.DEVNULL, env=asan_env)
        else:
            subprocess.run([a.exe or os.path.join(ORACLE, "_ref", "ref_driver"), fin, fout], check=True,
                           stdout=subprocess.DEVNULL)
        d = read_dump(fout)
        keep = {}
        for k, v in d.items():
            if name not in FULL and (k.startswith("ax_") or k.endswith("bx") or k == "dnm_sample"):
                continue
            keep[k] = v
        keep.update(in_xi=xi, in_ni=ni_n, in_xj=xj, in_nj=nj_n, in_q=q, in_box=box,
                    in_cfg=np.array([ni, nj, lmax, neq, int(same), int(asis), seed], np.int32))
        dst = os.path.join(a.out, f"ref_{name}.npz")
        np.savez_compressed(dst, **keep)
        print(f"{name}: {len(keep)} arrays -> {os.path.relpath(dst, ROOT)} ({os.path.getsize(dst)//1024} KiB)")
        os.remove(fin)
        os.remove(fout)


if __name__ == "__main__":
    sys.exit(main())
