"""Adjoints and position gradients of the matvec (old_fmm_bem_amd.adjoint):
central finite differences of the FMM energy, FP64 direct-sum gradients, the
transpose identity <u, K v> = <K^T u, v>, and the torch.autograd wrapper."""
import numpy as np
import pytest

from conftest import rel_l2
from oracle import oracle as O


def plummer(n, seed, rmax=10.0):
    """Plummer sphere (scale 1) truncated at r <= rmax, charges U[0,1)."""
    rng = np.random.default_rng(seed)
    out = np.zeros((0, 3))
    while len(out) < n:
        m = rng.random(2 * n)
        r = 1.0 / np.sqrt(np.maximum(m, 1e-12) ** (-2.0 / 3.0) - 1.0)
        v = rng.standard_normal((2 * n, 3))
        v /= np.linalg.norm(v, axis=1)[:, None]
        out = np.concatenate([out, (r[:, None] * v)[r <= rmax]])
    return out[:n].astype(np.float32), rng.random(n).astype(np.float32)


def test_position_gradient_formula_direct():
    """The identity the adjoint relies on, on the oracle's FP64 direct sum:
    dE/dx_k = q_k grad phi(x_k) for E = 1/2 q.G q (self pair excluded)."""
    rng = np.random.default_rng(3)
    x = rng.random((300, 3)).astype(np.float32)
    q = rng.random(300).astype(np.float32)

    def energy(xx):
        v, _ = O.direct(xx, xx, q, conv=O.EXACT)
        return 0.5 * float(q.astype(np.float64) @ v)

    _, g = O.direct(x, x, q, conv=O.EXACT)
    for k, a in [(0, 0), (17, 1), (123, 2), (299, 0)]:
        assert central_difference(energy, x, k, a) == pytest.approx(q[k] * g[k, a], rel=1e-6)


def central_difference(energy, x, k, a, h=1e-4):
    """dE/dx[k, a] from central differences at steps h and h/2 (float32
    coordinates; the step actually taken is used) and one Richardson step,
    which removes the h^2 term."""
    def d(step):
        xp, xm = x.copy(), x.copy()
        xp[k, a] = np.float32(x[k, a] + np.float32(step))
        xm[k, a] = np.float32(x[k, a] - np.float32(step))
        return (energy(xp) - energy(xm)) / (float(xp[k, a]) - float(xm[k, a]))

    return (4 * d(h / 2) - d(h)) / 3


def _away_from_faces(x, lmax, margin):
    f = x * (1 << lmax)
    return np.all(np.abs(f - np.round(f)) > margin * (1 << lmax), axis=1)


@pytest.mark.gpu
def test_energy_gradient_matches_finite_differences():
    """dE/dx and dE/dq of the FP64 FMM energy against central differences of
    the same FMM energy (h = 1e-4, 64 coordinates away from box faces), to
    1e-6."""
    import old_fmm_bem_amd as F
    from old_fmm_bem_amd.adjoint import energy_and_gradients

    n, lmax = 20000, 3
    rng = np.random.default_rng(40)
    x = rng.random((n, 3)).astype(np.float32)
    q = rng.random(n).astype(np.float32)
    f = F.FMM(p=12, lmax=lmax, precision="fp64", convention="exact", box=(0, 0, 0, 1))
    f.set_points(x)
    E, dq, dx = energy_and_gradients(f, q)
    ok = np.flatnonzero(_away_from_faces(x, lmax, 3e-4))
    picks = rng.choice(ok, 64, replace=False)
    axes = rng.integers(0, 3, 64)
    def energy(xx):
        return energy_and_gradients(f.set_points(xx), q)[0]

    fd, an = [], []
    for k, a in zip(picks, axes):
        fd.append(central_difference(energy, x, k, a))
        an.append(dx[k, a])
    fd, an = np.array(fd), np.array(an)
    print("max |fd - an| / max |an| =", np.abs(fd - an).max() / np.abs(an).max(), flush=True)
    assert np.abs(fd - an).max() <= 1e-6 * np.abs(an).max()
    # E is quadratic in q: the charge gradient K q from differences of E
    f.set_points(x)
    for k in picks[:8]:
        qp, qm = q.copy(), q.copy()
        qp[k] = np.float32(q[k] + np.float32(1e-3))
        qm[k] = np.float32(q[k] - np.float32(1e-3))
        Ep, _, _ = energy_and_gradients(f, qp)
        Em, _, _ = energy_and_gradients(f, qm)
        assert (Ep - Em) / (float(qp[k]) - float(qm[k])) == pytest.approx(dq[k], rel=1e-6)
    f.close()


@pytest.mark.gpu
def test_autograd_potential_plummer_vs_direct():
    """torch.autograd through the mixed-precision matvec on a Plummer sphere:
    q.grad and x.grad of L = w.phi against FP64 direct-sum formulas on 400
    sampled points."""
    import torch

    from old_fmm_bem_amd.adjoint import potential

    n = 60000
    x, q = plummer(n, 41)
    w = np.random.default_rng(42).random(n).astype(np.float32)
    xt = torch.from_numpy(x).cuda().requires_grad_(True)
    qt = torch.from_numpy(q).cuda().requires_grad_(True)
    wt = torch.from_numpy(w).cuda().to(torch.float64)
    phi = potential(xt, qt, p=10, lmax=5, precision="mixed", convention="exact")
    (wt * phi).sum().backward()
    idx = np.random.default_rng(43).choice(n, 400, replace=False)
    dvw, dgw = O.direct(x[idx], x, w, conv=O.EXACT)
    _, dgq = O.direct(x[idx], x, q, conv=O.EXACT)
    gq_ref = dvw
    gx_ref = w[idx, None].astype(np.float64) * dgq + q[idx, None].astype(np.float64) * dgw
    gq = qt.grad.detach().cpu().numpy()[idx]
    gx = xt.grad.detach().cpu().numpy()[idx]
    print("q.grad", rel_l2(gq, gq_ref), "x.grad", rel_l2(gx, gx_ref), flush=True)
    assert rel_l2(gq, gq_ref) < 1e-5
    assert rel_l2(gx, gx_ref) < 5e-5


@pytest.mark.gpu
@pytest.mark.parametrize("neq", [0, 1, 2])
def test_transpose_identity(neq):
    """<u, K v> = <K^T u, v> with K^T from adjoint() (Gni^T = -Gnj on one point
    set with one set of normals)."""
    import old_fmm_bem_amd as F
    from old_fmm_bem_amd.adjoint import adjoint

    rng = np.random.default_rng(44 + neq)
    n = 15000
    x = rng.random((n, 3)).astype(np.float32)
    nrm = rng.standard_normal((n, 3)).astype(np.float32)
    u = rng.standard_normal(n).astype(np.float32)
    v = rng.standard_normal(n).astype(np.float32)
    f = F.FMM(p=12, lmax=3, precision="fp64", convention="exact", box=(0, 0, 0, 1))
    f.set_points(x, x, ni_normals=nrm, nj_normals=nrm)
    Kv, _ = f.evaluate(v, neq=neq, gradient=False)
    KTu = adjoint(f, u, neq)
    lhs, rhs = float(u.astype(np.float64) @ Kv), float(KTu @ v.astype(np.float64))
    print(f"neq={neq} lhs={lhs} rhs={rhs}", flush=True)
    # normwise: a wrong sign or a missing transpose is off by ~|u||Kv|/sqrt(n)
    assert abs(lhs - rhs) <= 1e-6 * np.linalg.norm(u) * np.linalg.norm(Kv)
    f.close()


def _two_sets(seed, ni=9000, nj=11000):
    rng = np.random.default_rng(seed)
    xi = rng.random((ni, 3)).astype(np.float32)
    xj = (0.05 + 0.9 * rng.random((nj, 3))).astype(np.float32)
    ni_n = rng.standard_normal((ni, 3)).astype(np.float32)
    nj_n = rng.standard_normal((nj, 3)).astype(np.float32)
    q = (2 * rng.random(nj) - 1).astype(np.float32)
    w = (2 * rng.random(ni) - 1).astype(np.float32)
    return xi, xj, ni_n, nj_n, q, w


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("mixed", 3e-6)])
@pytest.mark.parametrize("neq", [0, 1, 2])
def test_transpose_distinct_sets(neq, precision, tol):
    """fmmx_evaluate_transpose with different target and source sets: K^T w
    against the oracle's direct sum with the roles exchanged (G symmetric;
    Gni^T = -Gnj carrying the target normals, Gnj^T = -Gni carrying the
    source normals) to the p = 12 truncation (2e-5; measured 2-5e-6), the
    identity <w, K q> = <K^T w, q> normwise to tol, and for G the gradient
    of the transposed field at the sources."""
    import old_fmm_bem_amd as F

    xi, xj, ni_n, nj_n, q, w = _two_sets(50 + neq)
    print(f"neq={neq} precision={precision}", flush=True)
    f = F.FMM(p=12, lmax=4, precision=precision, convention="exact", box=(0, 0, 0, 1))
    f.set_points(xi, xj, ni_normals=ni_n, nj_normals=nj_n)
    u, gu = f.transpose(w, neq=neq, gradient=(neq == 0))
    if neq == 0:
        ou, og = O.direct(xj, xi, w, conv=O.EXACT)
    elif neq == 1:
        ou, _ = O.direct(xj, xi, w, neq=2, conv=O.EXACT, nj_normals=ni_n)
        ou = -ou
    else:
        ou, _ = O.direct(xj, xi, w, neq=1, conv=O.EXACT, ni_normals=nj_n)
        ou = -ou
    e = rel_l2(u, ou)
    eg = rel_l2(gu, og) if neq == 0 else 0.0
    Kq, _ = f.evaluate(q, neq=neq, gradient=False)
    lhs, rhs = float(w.astype(np.float64) @ Kq), float(u @ q.astype(np.float64))
    ident = abs(lhs - rhs) / (np.linalg.norm(w) * np.linalg.norm(Kq))
    print(f"K^T w vs oracle {e:.2e}, gradient {eg:.2e}, identity {ident:.2e}", flush=True)
    assert e < 2e-5 and eg < 6e-5
    assert ident <= tol
    f.close()


@pytest.mark.gpu
def test_gradients_distinct_sets_vs_direct_and_differences():
    """fmmx_gradients for L = w . K q with different target and source sets
    (FP64, p = 12): dL/dq, dL/dx_i and dL/dx_j against the oracle's direct
    sums, and dL/dx_j against central differences of L computed by the FMM
    itself (h = 1e-4, one Richardson step) to 1e-6."""
    import old_fmm_bem_amd as F

    xi, xj, _, _, q, w = _two_sets(60, ni=6000, nj=7000)
    f = F.FMM(p=12, lmax=3, precision="fp64", convention="exact", box=(0, 0, 0, 1))
    f.set_points(xi, xj)
    dq, dxi, dxj = f.gradients(q, w)
    ov_w, og_w = O.direct(xj, xi, w, conv=O.EXACT)
    _, og_q = O.direct(xi, xj, q, conv=O.EXACT)
    w64, q64 = w.astype(np.float64)[:, None], q.astype(np.float64)[:, None]
    e = (rel_l2(dq, ov_w), rel_l2(dxi, w64 * og_q), rel_l2(dxj, q64 * og_w))
    print("dq %.2e dxi %.2e dxj %.2e" % e, flush=True)
    assert max(e) < 1e-5  # p = 12 truncation (measured 1.4-1.5e-6)

    def loss(xx):
        f.set_points(xi, xx)
        v, _ = f.evaluate(q, gradient=False)
        return float(w.astype(np.float64) @ v)

    rng = np.random.default_rng(61)
    ok = np.flatnonzero(_away_from_faces(xj, 3, 3e-4))
    picks, axes = rng.choice(ok, 24, replace=False), rng.integers(0, 3, 24)
    fd = np.array([central_difference(loss, xj, k, a) for k, a in zip(picks, axes)])
    an = np.array([dxj[k, a] for k, a in zip(picks, axes)])
    print("max |fd - an| / max |an| =", np.abs(fd - an).max() / np.abs(an).max(), flush=True)
    assert np.abs(fd - an).max() <= 1e-6 * np.abs(an).max()
    f.close()


@pytest.mark.gpu
def test_autograd_two_point_sets():
    """torch.autograd through potential(x, q, xj): gradients of L = w . phi
    reach q, the targets and the sources, and equal fmmx_gradients."""
    import torch

    import old_fmm_bem_amd as F
    from old_fmm_bem_amd.adjoint import potential

    xi, xj, _, _, q, w = _two_sets(62, ni=5000, nj=6000)
    xt = torch.from_numpy(xi).cuda().requires_grad_(True)
    yt = torch.from_numpy(xj).cuda().requires_grad_(True)
    qt = torch.from_numpy(q).cuda().requires_grad_(True)
    wt = torch.from_numpy(w).cuda().to(torch.float64)
    phi = potential(xt, qt, yt, p=10, lmax=3, precision="fp64", box=(0, 0, 0, 1))
    (wt * phi).sum().backward()
    f = F.FMM(p=10, lmax=3, precision="fp64", convention="exact", box=(0, 0, 0, 1))
    f.set_points(xi, xj)
    dq, dxi, dxj = f.gradients(q, w)
    assert rel_l2(qt.grad.cpu().numpy(), dq) < 1e-6
    assert rel_l2(xt.grad.cpu().numpy(), dxi) < 1e-6
    assert rel_l2(yt.grad.cpu().numpy(), dxj) < 1e-6
    f.close()
