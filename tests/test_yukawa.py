"""Yukawa kernel (the reference's neq 10, kernel/Helmholtz_*.cxx): operators
on the CPU against direct sums, and the GPU engine against the reference's
own outputs (tests/golden/ref_y*.npz) and the oracle's direct sum."""
import numpy as np
import pytest
from scipy.special import lpmv, spherical_in, spherical_kn

from conftest import golden, rel_l2, uniform_cube
import old_fmm_bem_amd as F
from oracle import oracle as O


# ---- host operators (no GPU) ------------------------------------------------
def _ibar(n, x):
    if x == 0:
        return 1.0 / np.prod(np.arange(1, 2 * n + 2, 2, dtype=float))
    return spherical_in(n, x) / x ** n


def _Y(n, m, v):
    from math import factorial
    r = np.linalg.norm(v)
    am = abs(m)
    fac = np.sqrt(factorial(n - am) / factorial(n + am))
    return fac * lpmv(am, n, v[2] / r if r > 0 else 1.0) * np.exp(1j * m * np.arctan2(v[1], v[0]))


def _p2m(p, lam, rb, pts, q, c):
    """scaled Yukawa multipole in the engine's real slots"""
    nr = p * (p + 1) // 2
    v = np.zeros(p * p)
    for n in range(p):
        for m in range(n + 1):
            s = 0
            for y, qq in zip(pts, q):
                d = y - c
                rho = np.linalg.norm(d)
                s += qq * _ibar(n, lam * rho) / _ibar(n, lam * rb) * (rho / rb) ** n * np.conj(_Y(n, m, d))
            v[n * (n + 1) // 2 + m] = s.real
            if m:
                v[nr + n * (n - 1) // 2 + m - 1] = s.imag
    return v


def _l2p(p, lam, rb, L, x, c):
    nr = p * (p + 1) // 2
    d = x - c
    r = np.linalg.norm(d)
    phi = 0.0
    for n in range(p):
        h = _ibar(n, lam * r) / _ibar(n, lam * rb) * (r / rb) ** n
        for m in range(n + 1):
            lc = L[n * (n + 1) // 2 + m] + (1j * L[nr + n * (n - 1) // 2 + m - 1] if m else 0)
            phi += (1 if m == 0 else 2) * (h * lc * _Y(n, m, d)).real
    return phi


@pytest.mark.parametrize("lam", [0.05, 0.9, 5.0])
@pytest.mark.parametrize("off", [(2, 0, 0), (2, -2, 1), (3, 3, 3)])
def test_yukawa_m2l_operator_vs_direct(lam, off):
    """P2M -> M2L operator -> L2P reproduces sum q k0(lam r) to the truncation
    error of p = 12 (the operator is the exact truncated re-expansion)."""
    rng = np.random.default_rng(7)
    c = -np.array(off, float)
    src = c + (rng.random((4, 3)) - 0.5) * 0.3
    q = rng.random(4)
    L = F.operator(12, "exact", 5, *off, rb=lam) @ _p2m(12, lam, 1.0, src, q, c)
    for x in (rng.random((4, 3)) - 0.5) * 0.3:
        ref = sum(qq * spherical_kn(0, lam * np.linalg.norm(x - y)) for y, qq in zip(src, q))
        assert abs(_l2p(12, lam, 1.0, L, x, np.zeros(3)) - ref) <= 1e-9 * abs(ref)


@pytest.mark.parametrize("lam", [0.05, 5.0])
def test_yukawa_m2l_operator_order_18(lam):
    """The reference's second order (mp = 18, kernel/precalch.cxx:185-261):
    the p = 18 operator reaches the truncation error of p = 18, far below
    what p = 12 gives for the same sources and targets."""
    rng = np.random.default_rng(9)
    off = (2, 1, 0)
    c = -np.array(off, float)
    src = c + (rng.random((3, 3)) - 0.5) * 0.5
    q = rng.random(3)
    xs = (rng.random((3, 3)) - 0.5) * 0.5
    err = {}
    for p in (12, 18):
        L = F.operator(p, "exact", 5, *off, rb=lam) @ _p2m(p, lam, 1.0, src, q, c)
        e = 0.0
        for x in xs:
            ref = sum(qq * spherical_kn(0, lam * np.linalg.norm(x - y)) for y, qq in zip(src, q))
            e = max(e, abs(_l2p(p, lam, 1.0, L, x, np.zeros(3)) - ref) / abs(ref))
        err[p] = e
    assert err[18] < 1e-10 and err[18] < 1e-2 * err[12], err


@pytest.mark.parametrize("oct_", [0, 3, 5, 6])
def test_yukawa_m2m_l2l_operators(oct_):
    """M2M(P2M about the child) ~ P2M about the parent (the Yukawa shift of a
    truncated multipole is not exact; its error is tiny); L2L preserves the
    local field at points of the child."""
    p, lam = 10, 0.9
    rng = np.random.default_rng(8 + oct_)
    cc = np.array([((oct_ >> 1) & 1) - 0.5, (oct_ & 1) - 0.5, ((oct_ >> 2) & 1) - 0.5]) * 0.5
    pts = cc + (rng.random((5, 3)) - 0.5) * 0.5
    q = rng.random(5)
    Mc = _p2m(p, lam, 0.5, pts, q, cc)
    Mp = _p2m(p, lam, 1.0, pts, q, np.zeros(3))
    A = F.operator(p, "exact", 3, oct_, rb=lam * 0.5)
    assert np.abs(A @ Mc - Mp).max() <= 1e-7 * np.abs(Mp).max()
    src = np.array([3.0, 1.0, -1.0])
    Lp = F.operator(p, "exact", 5, -3, -1, 1, rb=lam) @ _p2m(p, lam, 1.0, (src + 0.1)[None], np.ones(1), src)
    Lc = F.operator(p, "exact", 4, oct_, rb=lam * 0.5) @ Lp
    for x in cc + (rng.random((4, 3)) - 0.5) * 0.5:
        a, b = _l2p(p, lam, 1.0, Lp, x, np.zeros(3)), _l2p(p, lam, 0.5, Lc, x, cc)
        assert abs(a - b) <= 1e-10 * abs(a)


# ---- GPU engine ---------------------------------------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("case", ["y2", "y2l1"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_yukawa_vs_reference(case, precision):
    """The reference's Yukawa FMM (plane-wave M2L at mp=9) is accurate to about
    three digits against its own direct sum.  At the same order and
    conventions the engine (exact translations) lands closer to that direct
    sum and within the reference FMM's error of the reference FMM."""
    d = golden(case)
    ni, nj, lmax = (int(v) for v in d["in_cfg"][:3])
    lam = float(d["yukawa"][0])
    print(f"case={case} precision={precision} lam={lam}", flush=True)
    f = F.FMM(p=int(d["meta"][0]), lmax=lmax, precision=precision, convention="ref", box=tuple(d["in_box"]),
              lam=lam)
    f.set_points(d["in_xi"])
    v, g = f.evaluate(d["in_q"], neq=F.YUKAWA)
    e_ref = rel_l2(d["fmm_vi"], d["dir_vi"])
    e_us = rel_l2(v, d["dir_vi"])
    print(f"reference FMM vs direct {e_ref:.3e}, engine vs direct {e_us:.3e}", flush=True)
    assert e_us < 0.1 * e_ref
    assert rel_l2(v, d["fmm_vi"]) < 1.5 * e_ref
    # the GPU direct sum equals the reference's (float outputs)
    dv, _ = f.direct(d["in_xi"], d["in_xj"], d["in_q"], neq=F.YUKAWA)
    assert rel_l2(dv, d["dir_vi"]) < 1e-7
    f.close()


@pytest.mark.gpu
@pytest.mark.parametrize("lam", [0.5, 4.0, 30.0])
@pytest.mark.parametrize("precision,p,tol", [("fp64", 12, 5e-6), ("fp64", 8, 1e-4), ("mixed", 10, 2e-5)])
def test_yukawa_vs_direct(lam, precision, p, tol):
    """Unsoftened Yukawa FMM (self pair excluded) against the oracle's FP64
    direct sum, potential and gradient, on 600 sampled targets."""
    n = 30000
    x, q = uniform_cube(n, 60)
    print(f"lam={lam} precision={precision} p={p}", flush=True)
    f = F.FMM(p=p, lmax=4, precision=precision, convention="exact", box=(0, 0, 0, 1), lam=lam)
    f.set_points(x)
    v, g = f.evaluate(q, neq=F.YUKAWA)
    idx = np.random.default_rng(61).choice(n, 600, replace=False)
    dv, dg = O.direct_yukawa(x[idx], x, q, lam, conv=O.EXACT)
    ev, eg = rel_l2(v[idx], dv), rel_l2(g[idx], dg)
    print(f"potential {ev:.3e} gradient {eg:.3e}", flush=True)
    assert ev < tol and eg < 5 * tol
    # GPU direct == oracle direct
    gv, gg = f.direct(x[idx], x, q, neq=F.YUKAWA)
    assert rel_l2(gv, dv) < 1e-12 and rel_l2(gg, dg) < 1e-12
    f.close()


@pytest.mark.gpu
def test_yukawa_strong_screening_is_near_field_only():
    """lambda * (leaf box) beyond the far-field cut: the matvec is the near
    field alone, which then equals the direct sum to rounding."""
    n = 20000
    x, q = uniform_cube(n, 62)
    lam = 600.0  # leaf box 1/8: lambda rb = 75
    f = F.FMM(p=6, lmax=3, precision="fp64", convention="exact", box=(0, 0, 0, 1), lam=lam)
    f.set_points(x)
    v, g = f.evaluate(q, neq=F.YUKAWA)
    idx = np.arange(0, n, 40)
    dv, dg = O.direct_yukawa(x[idx], x, q, lam, conv=O.EXACT)
    assert rel_l2(v[idx], dv) < 1e-12 and rel_l2(g[idx], dg) < 1e-12
    f.close()


@pytest.mark.gpu
def test_yukawa_needs_lambda():
    f = F.FMM(p=6, lmax=3, precision="fp64", convention="exact", box=(0, 0, 0, 1))
    x, q = uniform_cube(1000, 63)
    f.set_points(x)
    with pytest.raises(F.FMMError):
        f.evaluate(q, neq=F.YUKAWA)
    with pytest.raises(F.FMMError):
        f.set_yukawa(-1.0)
    f.close()
