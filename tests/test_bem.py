"""Boundary-element layer (old_fmm_bem_amd.bem): quadrature and 6-node
elements of biobem/bem3d.cxx, the collocation operator of
biobem/fmm_matmult.h, GMRES and iterative refinement.  On the unit sphere the
singularity-subtracted double layer has the closed-form eigenvalues
1/2 - 1/(2(2n+1)) on degree-n spherical harmonics, which pins the whole
assembly (geometry, quadrature, normals, subtraction)."""
import math

import numpy as np
import pytest
from scipy.special import sph_harm_y

from conftest import rel_l2
from old_fmm_bem_amd import bem
from oracle import oracle as O


def sh_field(pts, nmax, seed):
    """Random real spherical-harmonic field of degree <= nmax at unit-sphere
    points, and its per-degree parts (to apply degree-wise eigenvalues)."""
    rng = np.random.default_rng(seed)
    th = np.arccos(np.clip(pts[:, 2] / np.linalg.norm(pts, axis=1), -1, 1))
    ph = np.arctan2(pts[:, 1], pts[:, 0])
    parts = []
    for n in range(nmax + 1):
        f = np.zeros(len(pts))
        for m in range(n + 1):
            y = sph_harm_y(n, m, th, ph)
            f += rng.standard_normal() * y.real
            if m:
                f += rng.standard_normal() * y.imag
        parts.append(f)
    return sum(parts), parts


def eig(n):
    return 0.5 - 1 / (2 * (2 * n + 1))


@pytest.mark.parametrize("ngl,degree", [(1, 1), (3, 2), (4, 3), (6, 4), (7, 5)])
def test_gauss_rules_exact_degree(ngl, degree):
    xi, eta, w = bem.gauss_trgl(ngl)
    assert w.sum() == pytest.approx(1.0, abs=1e-14)
    for a in range(degree + 1):
        for b in range(degree + 1 - a):
            exact = math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)
            assert 0.5 * np.sum(w * xi ** a * eta ** b) == pytest.approx(exact, rel=1e-12, abs=1e-15)


def test_shape_functions():
    rng = np.random.default_rng(1)
    X = rng.random((5, 6, 3))
    al, be, ga = 0.3 + 0.4 * rng.random((3, 5))
    nodes = [(0, 0), (1, 0), (0, 1), (al, 0 * al), (ga, 1 - ga), (0 * be, be)]
    for i, (xi, eta) in enumerate(nodes):
        ph, *_ = bem.interp_p(X, al, be, ga, xi + 0 * al, eta + 0 * al)
        assert np.allclose(ph, np.eye(6)[i][None, :], atol=1e-12)  # Kronecker at the nodes
    ph, dxi, det, vn, hsd = bem.interp_p(X, al, be, ga, 0.2, 0.3)
    assert np.allclose(ph.sum(1), 1)
    h = 1e-6
    p1, *_ = bem.interp_p(X, al, be, ga, 0.2 + h, 0.3)
    p0, *_ = bem.interp_p(X, al, be, ga, 0.2 - h, 0.3)
    assert np.allclose(np.einsum("ek,ekc->ec", (p1 - p0) / (2 * h), X), dxi, atol=1e-7)


def test_icosphere_geometry_converges():
    errs = []
    for s in (1, 2, 3, 4):
        m = bem.icosphere(s)
        assert m.nwe == 20 * 4 ** s and m.nwn == 40 * 4 ** s + 2
        assert np.all(np.sum(m.normals * m.nodes, 1) > 0.999)
        errs.append(abs(m.area.sum() - 4 * np.pi))
    assert all(b < a / 8 for a, b in zip(errs, errs[1:])), errs


@pytest.mark.parametrize("n", [1, 2, 4])
def test_subtracted_double_layer_eigenvalues(n):
    """sum_g w_g dG/dn_g (u(y_g) - u(x_l)) on the unit sphere, with the
    oracle's direct sum, against 1/2 - 1/(2(2n+1)) for Y_n^0."""
    m = bem.icosphere(3)
    u = sh_field(m.nodes, n, 5)[1][n]
    xi, xj, nj = (np.asarray(a, np.float32) for a in (m.nodes, m.qx, m.qn))
    s1, _ = O.direct(xi, xj, (m.qw * m.interpolate(u)).astype(np.float32), neq=O.NEQ_GNJ, conv=O.EXACT,
                     nj_normals=nj)
    k1, _ = O.direct(xi, xj, (-m.qw).astype(np.float32), neq=O.NEQ_GNJ, conv=O.EXACT, nj_normals=nj)
    y = s1 + k1 * u
    assert rel_l2(y, eig(n) * u) < 2e-4


def test_gmres_dense_system():
    """GMRES on a small dense nonsymmetric system, restarted, to 1e-12."""
    rng = np.random.default_rng(2)
    A = np.eye(300) * 3 + rng.standard_normal((300, 300)) / 30
    b = rng.standard_normal(300)
    x, hist, it = bem.gmres(lambda v: A @ v, b, rtol=1e-12, restart=20)
    assert np.linalg.norm(A @ x - b) <= 1e-11 * np.linalg.norm(b) and hist[-1] <= 1e-12


# ---- GPU ----------------------------------------------------------------------
def _direct_operator(m, e1, e2, u):
    xi, xj, nj = (np.asarray(a, np.float32) for a in (m.nodes, m.qx, m.qn))
    s1, _ = O.direct(xi, xj, (m.qw * m.interpolate(u)).astype(np.float32), neq=O.NEQ_GNJ, conv=O.EXACT,
                     nj_normals=nj)
    k1, _ = O.direct(xi, xj, (-m.qw).astype(np.float32), neq=O.NEQ_GNJ, conv=O.EXACT, nj_normals=nj)
    return (s1 + (k1 + (e1 + e2) / (e1 - e2)) * u) / e1


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [("fp64", 5e-6), ("mixed", 1e-5)])
def test_collocation_matvec_matches_direct(precision, tol):
    m = bem.icosphere(4)
    u = sh_field(m.nodes, 8, 6)[0]
    print(f"precision={precision} nodes={m.nwn} sources={len(m.qx)}", flush=True)
    op = bem.Collocation(m, e1=2.0, e2=80.0, p=12 if precision == "fp64" else 10, precision=precision)
    y = op.matvec(u)
    assert rel_l2(y, _direct_operator(m, 2.0, 80.0, u)) < tol
    op.close()


@pytest.mark.gpu
def test_gmres_sphere_against_closed_form():
    """Solve the biobem collocation system for a degree <= 8 harmonic
    right-hand side with GMRES on the mixed-precision FMM; the solution
    follows from the eigenvalues degree by degree."""
    e1, e2 = 2.0, 80.0
    m = bem.icosphere(4)
    b, parts = sh_field(m.nodes, 8, 7)
    diag = (e1 + e2) / (e1 - e2)
    exact = sum(p / ((eig(n) + diag) / e1) for n, p in enumerate(parts))
    op = bem.Collocation(m, e1=e1, e2=e2, p=10, precision="mixed")
    x, hist, it = bem.gmres(op.matvec, b, precond=op.precondition, rtol=1e-7, restart=30)
    print(f"iterations {it}, residual {hist[-1]:.2e}, error {rel_l2(x, exact):.2e}", flush=True)
    assert hist[-1] <= 1e-6 and it < 60
    assert rel_l2(x, exact) < 1e-4
    op.close()


@pytest.mark.gpu
def test_iterative_refinement_reaches_fp64_residual():
    """Mixed-precision inner GMRES (to 1e-5) inside FP64 residual sweeps:
    the residual of the FP64 operator drops to 1e-9 (its single-precision
    strength input bounds it near 1e-10)."""
    e1, e2 = 2.0, 80.0
    m = bem.icosphere(3)
    b, _ = sh_field(m.nodes, 6, 8)
    lo = bem.Collocation(m, e1=e1, e2=e2, p=10, precision="mixed")
    hi = bem.Collocation(m, e1=e1, e2=e2, p=12, precision="fp64")
    x, hist = bem.solve_refined(lo, hi, b, rtol=1e-9, inner_rtol=1e-5)
    print("refinement residuals", ["%.1e" % h for h in hist], flush=True)
    assert hist[-1] <= 1e-9 and len(hist) <= 5
    lo.close()
    hi.close()


@pytest.mark.gpu
def test_single_and_double_layer_on_sphere_points():
    """Single- and double-layer matvecs on the quadrature points of an
    icosphere (subdivision 5: 61,440 points, unit normals, a degree <= 8
    harmonic density): mixed precision vs the FP64 direct sum."""
    import old_fmm_bem_amd as F

    m = bem.icosphere(5)
    x = m.qx.astype(np.float32)
    nrm = (m.qx / np.linalg.norm(m.qx, axis=1)[:, None]).astype(np.float32)
    sig = sh_field(m.qx, 8, 9)[0]
    q = (m.qw * sig).astype(np.float32)
    f = F.FMM(p=10, precision="mixed", convention="exact")
    f.set_points(x, x, ni_normals=nrm, nj_normals=nrm)
    idx = np.random.default_rng(10).choice(len(x), 1000, replace=False)
    for neq, tol in ((F.G, 1e-5), (F.GNJ, 5e-5)):
        v, _ = f.evaluate(q, neq=neq, gradient=False)
        dv, _ = O.direct(x[idx], x, q, neq=neq, conv=O.EXACT, nj_normals=nrm)
        print(f"neq={neq} rel-L2 {rel_l2(v[idx], dv):.2e}", flush=True)
        assert rel_l2(v[idx], dv) < tol
    f.close()
