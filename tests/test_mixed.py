"""The mixed-precision engine against the FP64 engine on the same inputs, over
the configurations its FP32 kernels specialise on: every expansion order
(MFMA tiles of p^2 = 4 .. 144 rows, padded multipole layouts, the streaming
P2M/L2P forms), both conventions (reflection-canonical or per-level M2L
operators), sparse and deep trees (the one-launch M2L over 1 to 5 levels, empty
source levels), leaves small enough for two P2P lanes per target, and the edge
cases of the FP64 tests.  The FP64 engine runs separate kernels (vector-FMA
M2L, non-MFMA translations, one lane per target in P2P), so agreement to FP32
rounding checks the FP32 paths (measured 0.6-4.8e-7, bound 2e-6); the FP64
engine itself is checked against the oracle in test_gpu.py."""
import numpy as np
import pytest

from conftest import rel_l2, uniform_cube

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import old_fmm_bem_amd as F  # noqa: E402


def _both(p, lmax, conv, x, q, xj=None, neq=F.G, nrm=None, **kw):
    out = []
    for prec in ("mixed", "fp64"):
        f = F.FMM(p=p, lmax=lmax, precision=prec, convention=conv, box=(0, 0, 0, 1), **kw)
        f.set_points(x, xj, ni_normals=nrm, nj_normals=nrm)
        out.append(f.evaluate(q, neq=neq))
        f.close()
    return out


@pytest.mark.parametrize("conv", ["exact", "ref"])
@pytest.mark.parametrize("p", [2, 3, 5, 7, 9, 11, 12])
def test_mixed_matches_fp64_every_order(p, conv):
    x, q = uniform_cube(20000, 100 + p)
    nrm = np.random.default_rng(p).standard_normal((20000, 3)).astype(np.float32)
    for neq in (F.G, F.GNJ):
        print(f"p={p} conv={conv} neq={neq}", flush=True)
        (vm, gm), (vd, gd) = _both(p, 4, conv, x, q, neq=neq, nrm=nrm)
        ev = rel_l2(vm, vd)
        print(f"  potential {ev:.2e}", flush=True)
        assert ev < 2e-6
        if neq == F.G:
            eg = rel_l2(gm, gd)
            print(f"  gradient {eg:.2e}", flush=True)
            assert eg < 2e-6


@pytest.mark.parametrize("lmax", [2, 3, 5, 6])
def test_mixed_matches_fp64_sparse_deep_trees(lmax):
    """Points on a sphere: most boxes empty, upper levels with few tiles,
    levels where a class has no boxes."""
    rng = np.random.default_rng(40 + lmax)
    n = 60000
    s = rng.standard_normal((n, 3))
    x = (0.5 + 0.45 * s / np.linalg.norm(s, axis=1, keepdims=True)).astype(np.float32)
    q = (rng.random(n) * 2 - 1).astype(np.float32)
    print(f"lmax={lmax}", flush=True)
    (vm, gm), (vd, gd) = _both(8, lmax, "exact", x, q)
    ev, eg = rel_l2(vm, vd), rel_l2(gm, gd)
    print(f"  potential {ev:.2e} gradient {eg:.2e}", flush=True)
    assert ev < 2e-6 and eg < 2e-6


def test_mixed_edge_cases():
    rng = np.random.default_rng(11)
    xj, q = uniform_cube(3000, 51)
    # one target, disjoint target and source sets
    xi = np.array([[0.3, 0.4, 0.5]], np.float32)
    (vm, _), (vd, _) = _both(6, 3, "exact", xi, q, xj=xj)
    assert rel_l2(vm, vd) < 2e-6
    xi2, _ = uniform_cube(2000, 52)
    (vm, gm), (vd, gd) = _both(10, 4, "exact", xi2, q, xj=xj)
    assert rel_l2(vm, vd) < 2e-6 and rel_l2(gm, gd) < 2e-6
    # every point in one leaf box: near field only
    xc = (0.01 + 0.02 * rng.random((300, 3))).astype(np.float32)
    qc = rng.random(300).astype(np.float32)
    (vm, _), (vd, _) = _both(6, 3, "exact", xc, qc)
    assert rel_l2(vm, vd) < 1e-6
    # leaves of ~100 points (wide P2P workgroups) at lmax 2
    xw, qw = uniform_cube(8000, 53)
    (vm, gm), (vd, gd) = _both(10, 2, "exact", xw, qw)
    assert rel_l2(vm, vd) < 2e-6 and rel_l2(gm, gd) < 2e-6
    # empty source set
    f = F.FMM(p=6, lmax=3, precision="mixed", convention="exact", box=(0, 0, 0, 1))
    f.set_points(xi, np.zeros((0, 3), np.float32))
    v, g = f.evaluate(np.zeros(0, np.float32))
    assert v.shape == (1,) and v[0] == 0 and np.all(g == 0)
    f.close()


SPLIT_SHAPES = [(2, 5, 8), (2, 5, 4), (4, 5, 4), (1, 4, 8), (1, 3, 8), (1, 3, 4), (2, 3, 4)]


def _with_shape(f, shape, lmax):
    for lev in range(2, lmax + 1):
        f.set_m2l_variant(lev, *shape, 0)


@pytest.mark.parametrize("p", [6, 8, 10, 12])
def test_split_fp16_m2l_matches_fp32(p):
    """The error-compensated FP16 M2L (staging 3, 4: hi/lo operands, three
    products per pair on v_mfma_f32_16x16x32_f16; 5: on 32x32x16, p <= 10,
    32 or 64 targets per wave)
    against the FP32 MFMA M2L
    and the FP64 engine on the same tree, every launch shape: the split form
    carries 2^-22 relative operand error against FP32's 2^-24, so its
    difference from the FP64 engine stays at FP32 rounding (bound 2e-6, the
    FP32 engine's own) and from the FP32 M2L well below the p truncation."""
    x, q = uniform_cube(60000, 300 + p)
    lmax = 4
    res = {}
    for prec, shape in [("fp64", None), ("mixed", (1, 0, 8))] + [("mixed", s) for s in SPLIT_SHAPES]:
        f = F.FMM(p=p, lmax=lmax, precision=prec, convention="exact", box=(0, 0, 0, 1))
        f.set_points(x)
        if shape:
            _with_shape(f, shape, lmax)
        v, g = f.evaluate(q)
        eff = f.m2l_variant(lmax) if shape else None
        f.close()
        res[prec, shape] = (v, g, eff)
    vd, gd, _ = res["fp64", None]
    v32, g32, _ = res["mixed", (1, 0, 8)]
    for s in SPLIT_SHAPES:
        v, g, eff = res["mixed", s]
        assert eff[1] in (3, 4, 5), eff  # a split kernel ran
        ev, eg = rel_l2(v, vd), rel_l2(g, gd)
        e32 = rel_l2(v, v32)
        print(f"p={p} shape={s} -> {eff}: vs FP64 {ev:.2e} / {eg:.2e}, vs FP32 M2L {e32:.2e}", flush=True)
        assert ev < 2e-6 and eg < 2e-6
        assert e32 < 2e-6


def test_split_fp16_m2l_scale_invariant():
    """Charges scaled by 2^-40 and 2^40 (far outside FP16's range unscaled):
    the per-level power-of-two scaling of the split multipoles gives the same
    relative result."""
    x, q = uniform_cube(30000, 77)
    for shape in ((2, 5, 8), (1, 4, 8)):
        out = []
        for s in (1.0, 2.0 ** -40, 2.0 ** 40):
            f = F.FMM(p=10, lmax=4, precision="mixed", convention="exact", box=(0, 0, 0, 1))
            f.set_points(x)
            _with_shape(f, shape, 4)
            v, _ = f.evaluate((q * np.float32(s)).astype(np.float32))
            f.close()
            out.append(v / s)
        assert rel_l2(out[1], out[0]) < 1e-12 and rel_l2(out[2], out[0]) < 1e-12
