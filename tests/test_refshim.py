"""libfmmx_ref.so: the reference's own C++ entry points over the caller's
globals.  tests/refshim/ref_caller is a caller written against the reference
interface (the test/test.cxx:98-146 sequence) and linked only against the
shim; its outputs are checked against what the compiled reference wrote for
the same inputs (tests/golden/)."""
import os
import struct
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN_CASES, ROOT, golden, rel_l2

CALLER = os.path.join(ROOT, "tests", "refshim", "ref_caller")


def _write_input(path, d):
    ni, nj, lmax, neq = (int(v) for v in d["in_cfg"][:4])
    with open(path, "wb") as f:
        f.write(struct.pack("<5i", ni, nj, lmax, neq, 0))
        f.write(np.asarray(d["in_box"], np.float32).tobytes())
        for a in (d["in_xi"], d["in_ni"], d["in_xj"], d["in_nj"]):
            for k in range(3):
                f.write(np.ascontiguousarray(a[:, k], dtype=np.float32).tobytes())
        f.write(np.asarray(d["in_q"], np.float32).tobytes())
        if neq == 10:
            f.write(np.float32(d["yukawa"][0]).tobytes())


def _read_output(path):
    out, data, p = {}, open(path, "rb").read(), 0
    while p < len(data):
        name = data[p:p + 32].split(b"\0", 1)[0].decode()
        typ, n = struct.unpack_from("<2i", data, p + 32)
        p += 40
        dt = {0: np.int32, 1: np.float32, 2: np.float64}[typ]
        es = np.dtype(dt).itemsize
        out[name] = np.frombuffer(data[p:p + es * n], dtype=dt).copy()
        p += es * n
    return out


def _run(d, tmp_path, env=None, exe=CALLER, args=()):
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    _write_input(fin, d)
    r = subprocess.run([exe, fin, fout, *args], capture_output=True, text=True, timeout=120,
                       env=dict(os.environ, **(env or {})))
    return r, (_read_output(fout) if r.returncode == 0 else None)


def _need_caller():
    if not os.path.exists(CALLER):
        pytest.skip("tests/refshim/ref_caller not built (build() builds it)")


def test_caller_without_device_reports_and_exits(tmp_path):
    """Errors follow the reference's CUDA error path: a message and exit(1),
    never an abort.  Runs where no HIP device is visible."""
    _need_caller()
    r, _ = _run(golden("g2"), tmp_path, env={"HIP_VISIBLE_DEVICES": "-1", "ROCR_VISIBLE_DEVICES": "-1"})
    assert r.returncode == 1, (r.returncode, r.stderr)
    assert "fmmx_ref: " in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["y2", "y2l1"])
def test_reference_caller_yukawa(case, tmp_path):
    """fmm(ni, nj, 10) and Helmholtz_G_dir through the shim (test/testh.cxx
    sequence): permutations bit for bit, the direct sum as the reference's,
    the matvec closer to it than the reference's own three-digit FMM."""
    _need_caller()
    d = golden(case)
    print(f"case={case}", flush=True)
    r, o = _run(d, tmp_path)
    assert r.returncode == 0, r.stderr
    assert np.array_equal(o["nbi"], d["nbi"]) and np.array_equal(o["nbj"], d["nbj"])
    e_ref = rel_l2(d["fmm_vi"], d["dir_vi"])
    assert rel_l2(o["fmm_vi"], d["dir_vi"]) < 0.1 * e_ref
    assert rel_l2(o["fmm2_vi_sorted"], 2 * o["fmm_vi_sorted"]) < 1e-6
    assert rel_l2(o["dir_vi"], d["dir_vi"]) < 1e-7


def _check_outputs(o, d, neq):
    """The matvec outputs of a caller run against what the reference wrote."""
    assert rel_l2(o["fmm_vi_sorted"], d["fmm_vi_sorted"]) < 2e-6
    assert rel_l2(o["fmm_vi"], d["fmm_vi"]) < 2e-6
    if neq == 0:
        g = np.stack([o["fmm_gxi"], o["fmm_gyi"], o["fmm_gzi"]], 1)
        assert rel_l2(g, np.stack([d["fmm_gxi"], d["fmm_gyi"], d["fmm_gzi"]], 1)) < 2e-6
    # second matvec on the cached tree, strengths doubled
    assert rel_l2(o["fmm2_vi_sorted"], 2 * o["fmm_vi_sorted"]) < 1e-6
    assert np.array_equal(o["nbi"], d["nbi"]) and np.array_equal(o["nbj"], d["nbj"])


@pytest.mark.gpu
@pytest.mark.parametrize("case", GOLDEN_CASES)
def test_operator_symbols_match_reference(case, tmp_path):
    """The operator plugin (the symbols fmm/fmm.cxx:7-27 calls) sequenced as
    fmm/fmm.cxx:38-150 does, phase by phase, against the reference's own
    dumps: stacked box keys bit for bit, far lists bit for bit (level 2: the
    neib-1 list), multipoles after P2M and after the upward pass, locals of
    every level after its M2L (FP64 engine, REF convention: 1e-10), near
    field and outputs (float, 2e-6)."""
    _need_caller()
    d = golden(case)
    ni, nj, lmax, neq = (int(v) for v in d["in_cfg"][:4])
    print(f"case={case} neq={neq} lmax={lmax}", flush=True)
    r, o = _run(d, tmp_path, args=("plugin",))
    assert r.returncode == 0, r.stderr
    # level offsets of the stacked lists: entries 1..lmax-1 (entry 0, level 1,
    # is never written by the reference, fmm/boxm2mi.cxx:8)
    for k in ("nlbi", "nlbj"):
        assert np.array_equal(o[k][1:lmax], d[k][1:lmax]), k
    for k in ("all_nfi", "all_nfj"):
        assert np.array_equal(o[k], d[k]), k
    for lev in range(2, lmax + 1):
        assert np.array_equal(o[f"far_nij_{lev}"], d[f"far_nij_{lev}"]), lev
        assert np.array_equal(o[f"far_neij_{lev}"], d[f"far_neij_{lev}"]), lev
    assert rel_l2(o["p2p_vi"], d["p2p_vi"]) < 2e-6
    if "p2m_bx" in d:
        for k in ("p2m_bx", "up_bx"):
            assert np.abs(o[k] - d[k]).max() <= 1e-11 * np.abs(d[k]).max(), k
        for lev in range(2, lmax + 1):
            ref = d[f"ax_{lev}"]
            assert np.abs(o[f"ax_{lev}"] - ref).max() <= 1e-10 * np.abs(ref).max(), lev
    _check_outputs(o, d, neq)


REF_FMM = os.path.join(ROOT, "oracle", "_ref", "ref_caller_fmm")
REF_KERNELS = os.path.join(ROOT, "oracle", "_ref", "ref_caller_kernels")


@pytest.mark.gpu
@pytest.mark.parametrize("exe", [REF_FMM, REF_KERNELS], ids=["fmm.o", "fmm.o+tree"])
@pytest.mark.parametrize("case", ["g3", "gni3", "gnj3", "g2", "gedge4"])
def test_reference_fmm_object_with_shim_kernels(exe, case, tmp_path):
    """The reference's fmm/fmm.cxx compiled from its sources (oracle/Makefile.ref
    "callers", built where the reference is present) linked against the
    shim: with the shim's box symbols (fmm.o), and with all of the reference's
    tree, list and set-up objects so that only the kernels come from the shim
    (fmm.o+tree, biobem/Makefile:25-37).  Its own level-2 call asks for the
    list neib 3, which fmm/ijbox.cxx never builds; the shim's m2l uses the
    full level-2 list, so the outputs are those of the golden driver."""
    if not os.path.exists(exe):
        pytest.skip(f"{os.path.relpath(exe, ROOT)} not built (needs the reference sources)")
    d = golden(case)
    ni, nj, lmax, neq = (int(v) for v in d["in_cfg"][:4])
    print(f"exe={os.path.basename(exe)} case={case} neq={neq} lmax={lmax}", flush=True)
    r, o = _run(d, tmp_path, exe=exe)
    assert r.returncode == 0, r.stderr
    _check_outputs(o, d, neq)
    assert rel_l2(o["dir_vi"], d["dir_vi"]) < 1e-7


@pytest.mark.gpu
@pytest.mark.parametrize("case", GOLDEN_CASES)
def test_reference_caller_matches_reference(case, tmp_path):
    _need_caller()
    d = golden(case)
    ni, nj, lmax, neq = (int(v) for v in d["in_cfg"][:4])
    print(f"case={case} neq={neq} lmax={lmax}", flush=True)
    r, o = _run(d, tmp_path)
    assert r.returncode == 0, r.stderr
    # sort permutations and box counts: bit for bit (sort/sorti.cxx:21-23,
    # fmm/boxallocate.cxx:88-89)
    assert np.array_equal(o["nbi"], d["nbi"]) and np.array_equal(o["nbj"], d["nbj"])
    nbmax, nbne, nbnet, lm = o["counts"]
    assert (nbmax, lm) == (1 << 3 * lmax, lmax)
    assert nbne == max(len(d["leaf_nfi"]), len(d["leaf_nfj"]))
    assert nbnet == max(len(d["all_nfi"]), len(d["all_nfj"]))
    assert o["nek_boxes"][0] == len(np.union1d(d["leaf_nfi"], d["leaf_nfj"]))
    # matvec outputs, sorted and caller order, as the reference wrote them
    assert rel_l2(o["fmm_vi_sorted"], d["fmm_vi_sorted"]) < 2e-6
    assert rel_l2(o["fmm_vi"], d["fmm_vi"]) < 2e-6
    if neq == 0:
        g = np.stack([o["fmm_gxi"], o["fmm_gyi"], o["fmm_gzi"]], 1)
        assert rel_l2(g, np.stack([d["fmm_gxi"], d["fmm_gyi"], d["fmm_gzi"]], 1)) < 2e-6
    # second matvec on the cached tree, strengths doubled
    assert rel_l2(o["fmm2_vi_sorted"], 2 * o["fmm_vi_sorted"]) < 1e-6
    # unsortj restored the caller's arrays exactly
    assert np.array_equal(o["xj_restored"], d["in_xj"][:, 0]) and np.array_equal(o["vj_restored"], d["in_q"])
    assert o["tfmm_total"][0] > 0
    # direct sums (kernel/G_dir.cxx, Gni_dir.cxx, Gnj_dir.cxx)
    assert rel_l2(o["dir_vi"], d["dir_vi"]) < 1e-7
    if neq == 0:
        g = np.stack([o["dir_gxi"], o["dir_gyi"], o["dir_gzi"]], 1)
        assert rel_l2(g, np.stack([d["dir_gxi"], d["dir_gyi"], d["dir_gzi"]], 1)) < 1e-7
