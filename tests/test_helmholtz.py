"""Oscillatory Helmholtz kernel exp(i k r) / (4 pi r) with complex charges
(FMMX_HELMHOLTZ).  The reference has no such kernel; everything here is
pinned by direct sums: the host operators against scipy's spherical Bessel
functions, the GPU engine against the oracle's FP64 direct sum."""
import numpy as np
import pytest
from scipy.special import lpmv, spherical_jn

from conftest import rel_l2, uniform_cube
import old_fmm_bem_amd as F


def _df(n):  # (2n+1)!!
    return float(np.prod(np.arange(1, 2 * n + 2, 2, dtype=float)))


def _jhat(n, x):
    return 1.0 if x == 0 else spherical_jn(n, x) * _df(n) / x ** n


def _Y(n, m, v):
    from math import factorial
    r = np.linalg.norm(v)
    am = abs(m)
    fac = np.sqrt(factorial(n - am) / factorial(n + am))
    return fac * lpmv(am, n, v[2] / r if r > 0 else 1.0) * np.exp(1j * m * np.arctan2(v[1], v[0]))


def _slots(p, X):
    """all-m coefficients of a field with X_n^-m = conj(X_n^m) -> real slots"""
    nr = p * (p + 1) // 2
    v = np.zeros(p * p)
    for n in range(p):
        for m in range(n + 1):
            v[n * (n + 1) // 2 + m] = X[(n, m)].real
            if m:
                v[nr + n * (n - 1) // 2 + m - 1] = X[(n, m)].imag
    return v


def _p2m(p, k, rb, pts, q, c):
    """split multipole (Re-charge part, Im-charge part) about c"""
    out = []
    for qq in (q.real, q.imag):
        X = {}
        for n in range(p):
            for m in range(n + 1):
                s = 0
                for y, qj in zip(pts, qq):
                    d = y - c
                    rho = np.linalg.norm(d)
                    s += qj * _jhat(n, k * rho) * (rho / rb) ** n * np.conj(_Y(n, m, d))
                X[(n, m)] = s
        out.append(_slots(p, X))
    return np.concatenate(out)


def _l2p(p, k, rb, L, x, c):
    """complex potential of a split local expansion about c at x"""
    nr = p * (p + 1) // 2
    d = x - c
    r = np.linalg.norm(d)
    val = []
    for part in (L[:p * p], L[p * p:]):
        phi = 0.0
        for n in range(p):
            h = _jhat(n, k * r) * (r / rb) ** n
            for m in range(n + 1):
                lc = part[n * (n + 1) // 2 + m] + (1j * part[nr + n * (n - 1) // 2 + m - 1] if m else 0)
                phi += (1 if m == 0 else 2) * (h * lc * _Y(n, m, d)).real
        val.append(phi / (4 * np.pi * rb))
    return val[0] + 1j * val[1]


def _direct(x, pts, q, k):
    r = np.linalg.norm(pts - x, axis=1)
    return np.sum(q * np.exp(1j * k * r) / (4 * np.pi * r))


@pytest.mark.parametrize("kappa", [0.05, 1.5, 3.0])
@pytest.mark.parametrize("off", [(2, 0, 0), (2, -2, 1), (3, 3, 3)])
def test_helmholtz_m2l_operator_vs_direct(kappa, off):
    """P2M -> M2L operator -> L2P reproduces sum q exp(ikr)/(4 pi r), complex
    charges, to the truncation error of p = 12."""
    p = 12
    rng = np.random.default_rng(3)
    c = -np.array(off, float)
    src = c + (rng.random((4, 3)) - 0.5) * 0.3
    q = rng.random(4) - 0.5 + 1j * (rng.random(4) - 0.5)
    L = F.operator(p, "exact", 8, *off, rb=kappa) @ _p2m(p, kappa, 1.0, src, q, c)
    for x in (rng.random((3, 3)) - 0.5) * 0.3:
        ref = _direct(x, src, q, kappa)
        assert abs(_l2p(p, kappa, 1.0, L, x, np.zeros(3)) - ref) <= 1e-9 * abs(ref)


@pytest.mark.parametrize("oct_", [0, 3, 5, 6])
def test_helmholtz_m2m_l2l_operators(oct_):
    """M2M(P2M about the child) ~ P2M about the parent; L2L preserves the
    local field at points of the child (kappa = k * child box size)."""
    p, kc = 10, 0.7
    rng = np.random.default_rng(20 + oct_)
    cc = np.array([((oct_ >> 1) & 1) - 0.5, (oct_ & 1) - 0.5, ((oct_ >> 2) & 1) - 0.5]) * 0.5
    pts = cc + (rng.random((5, 3)) - 0.5) * 0.5
    q = rng.random(5) - 0.5 + 1j * (rng.random(5) - 0.5)
    k = 2 * kc  # parent box size 1
    Mc = _p2m(p, k, 0.5, pts, q, cc)
    Mp = _p2m(p, k, 1.0, pts, q, np.zeros(3))
    A = F.operator(p, "exact", 6, oct_, rb=kc)
    # the parent's degree-n coefficients also receive the child's degrees >= p
    # that truncation drops: a small error at the top degrees
    assert np.abs(A @ Mc - Mp).max() <= 3e-7 * np.abs(Mp).max()
    src = np.array([3.0, 1.0, -1.0])
    ps = (src + 0.1)[None]
    Lp = F.operator(p, "exact", 8, -3, -1, 1, rb=k) @ _p2m(p, k, 1.0, ps, np.array([1 + 0.5j]), src)
    Lc = F.operator(p, "exact", 7, oct_, rb=kc) @ Lp
    for x in cc + (rng.random((4, 3)) - 0.5) * 0.5:
        a, b = _l2p(p, k, 1.0, Lp, x, np.zeros(3)), _l2p(p, k, 0.5, Lc, x, cc)
        assert abs(a - b) <= 1e-10 * abs(a)
        assert abs(a - _direct(x, ps, np.array([1 + 0.5j]), k)) <= 1e-6 * abs(a)


def test_helmholtz_operator_small_k_is_laplace():
    """As k -> 0 the Re part of the M2L operator on the Re part is the
    Laplace operator in the same scaled basis."""
    p = 8
    A = F.operator(p, "exact", 8, 2, -1, 3, rb=1e-6)
    B = F.operator(p, "exact", 2, 2, -1, 3)
    P2 = p * p
    assert np.abs(A[:P2, :P2] - B).max() <= 1e-9 * np.abs(B).max()
    assert np.abs(A[P2:, P2:] - B).max() <= 1e-9 * np.abs(B).max()


def test_helmholtz_operator_arguments():
    A = np.zeros((200, 200))
    assert F.lib().fmmx_operator(10, 1, 8, 1, 0, 0, 1.0, A.ctypes.data) == 1  # adjacent offset
    assert F.lib().fmmx_operator(10, 1, 8, 2, 0, 0, 9.0, A.ctypes.data) == 1  # k rb too large
    assert F.lib().fmmx_operator(10, 1, 6, 8, 0, 0, 1.0, A.ctypes.data) == 1  # octant
