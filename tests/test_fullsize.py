"""The configurations of the matvec at their stated sizes, on the GPU, each
against an FP64 direct sum on sampled targets (the GPU direct kernel, itself
checked against the oracle's direct sum on a subset):

  cube     N = 1,000,000 uniform, p = 10, mixed and FP64 (the headline);
  C2       245,760 Gauss points of an 81,920-triangle icosphere, G and Gnj;
  C3       1,000,000 points on the unit sphere, Helmholtz k = 5, p = 12;
  C4       1,000,000 Plummer points, energy gradients of the mixed engine,
           and central differences of the energy (h = 1e-4, 64 coordinates);
  C5       256 ellipsoids, 11.6M panels, batched GMRES solve, p = 8, mixed.
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no HIP device", allow_module_level=True)

import old_fmm_bem_amd as F  # noqa: E402
from old_fmm_bem_amd import workloads as W  # noqa: E402
from oracle import oracle as O  # noqa: E402


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _rel(a, b):
    return float(torch.linalg.norm(a - b) / torch.linalg.norm(b))


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_cube_1m_p10(precision):
    """N = 1M uniform, p = 10, lmax 5: potential rel-L2 <= 1e-5 and gradient
    <= 3e-5 against the FP64 direct sum on 1,000 targets; on 100 of them the
    GPU direct sum equals the oracle's to 1e-12."""
    x, q = W.cube(1_000_000, 0)
    xd, qd = _dev(x), _dev(q)
    idx = np.random.default_rng(1).choice(len(x), 1000, replace=False)
    it = torch.from_numpy(idx).cuda()
    print(f"cube N=1M p=10 {precision}", flush=True)
    f = F.FMM(p=10, lmax=5, precision=precision, convention="exact", box=(0, 0, 0, 1))
    f.set_points(xd)
    v, g = f.evaluate(qd)
    dv, dg = f.direct(xd[it], xd, qd)
    ev, eg = _rel(v[it], dv), _rel(g[it], dg)
    print(f"potential {ev:.3e} gradient {eg:.3e}", flush=True)
    assert ev <= 1e-5 and eg <= 3e-5
    ov, og = O.direct(x[idx[:100]], x, q, conv=O.EXACT)
    assert rel_l2(dv[:100].cpu().numpy(), ov) < 1e-12 and rel_l2(dg[:100].cpu().numpy(), og) < 1e-12
    f.close()


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_bem_sphere_c2(precision):
    """C2: single layer (G) and double layer (Gnj) on the 245,760 Gauss points,
    p = 10, against the oracle's FP64 direct sum on 1,000 targets.  The double
    layer of so fine a surface sampling carries the p = 10 truncation
    (8.8e-5 in both precisions); mixed precision adds less than 2% to either
    error."""
    x, nrm, q = W.bem_sphere(6, 0)
    assert len(x) == 245_760
    idx = np.random.default_rng(1).choice(len(x), 1000, replace=False)
    it = torch.from_numpy(idx).cuda()
    kernels = (("G", F.G), ("Gnj", F.GNJ))
    ref = {name: O.direct(x[idx], x, q, neq=neq, conv=O.EXACT, nj_normals=nrm)[0] for name, neq in kernels}
    errs = {}
    for prec in ("fp64", precision):
        f = F.FMM(p=10, lmax=5, precision=prec, convention="exact")
        f.set_points(_dev(x), ni_normals=_dev(nrm), nj_normals=_dev(nrm))
        for name, neq in kernels:
            v, _ = f.evaluate(_dev(q), neq=neq, gradient=False)
            errs[prec, name] = rel_l2(v[it].cpu().numpy(), ref[name])
        f.close()
    print({f"{k[0]} {k[1]}": f"{e:.3e}" for k, e in errs.items()}, flush=True)
    assert errs["fp64", "G"] < 5e-6 and errs["fp64", "Gnj"] < 2e-4
    assert errs[precision, "G"] <= 1.02 * errs["fp64", "G"] + 1e-7
    assert errs[precision, "Gnj"] <= 1.02 * errs["fp64", "Gnj"] + 1e-7


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_helmholtz_sphere_c3(precision):
    """C3: 1M points on the unit sphere, k = 5 (k D = 10), complex charges,
    p = 12, lmax 5: rel-L2 <= 2e-6 against the GPU FP64 direct sum on 2,000
    targets, which equals the oracle's on 200 of them to 1e-12."""
    x, q = W.sphere(1_000_000, 0)
    xd, qd = _dev(x), _dev(q)
    idx = np.random.default_rng(1).choice(len(x), 2000, replace=False)
    it = torch.from_numpy(idx).cuda()
    print(f"helmholtz sphere N=1M k=5 p=12 {precision}", flush=True)
    f = F.FMM(p=12, lmax=5, precision=precision, convention="exact", k=5.0)
    f.set_points(xd)
    v, _ = f.evaluate(qd, neq=F.HELMHOLTZ)
    dv, _ = f.direct(xd[it], xd, qd, neq=F.HELMHOLTZ)
    e = _rel(v[it], dv)
    print(f"rel-L2 {e:.3e}", flush=True)
    assert e <= 2e-6
    ov = O.direct_helmholtz(x[idx[:200]], x, q, 5.0)
    assert rel_l2(dv[:200].cpu().numpy(), ov) < 1e-12
    f.close()


def _pair_difference_fd(f, xd, qd, k, a, h):
    """Central difference of the exact energy E = 1/2 sum_{i != j} q_i q_j G
    in coordinate (k, a), Richardson-extrapolated from steps h and h/2.  Only
    the pairs holding point k change, so E(x + s e) - E(x - s e) =
    q_k (phi_k(x_k + s e) - phi_k(x_k - s e)) with phi_k the potential of all
    other charges: two FP64 direct sums at displaced targets (charge k set to
    zero) whose difference keeps its digits, unlike a difference of two
    full energies (|E| ~ 1e10 against a difference ~ 1e1)."""
    qk = float(qd[k])
    qz = qd.clone()
    qz[k] = 0
    xk = xd[k].double().cpu().numpy()
    pts, steps = [], []
    for s in (h, h / 2):
        p, m = xk.copy(), xk.copy()
        p[a] = np.float32(xk[a] + np.float32(s))
        m[a] = np.float32(xk[a] - np.float32(s))
        pts += [p, m]
        steps.append(float(np.float32(p[a])) - float(np.float32(m[a])))
    phi, _ = f.direct(_dev(np.array(pts, np.float32)), xd, qz)
    phi = phi.cpu().numpy()
    d1 = qk * (phi[0] - phi[1]) / steps[0]
    d2 = qk * (phi[2] - phi[3]) / steps[1]
    return (4 * d2 - d1) / 3


def test_adjoint_plummer_c4():
    """C4: E = 1/2 q.G q on 1M Plummer points (r <= 10), q in [0,1], p = 10:
    the mixed engine's dE/dq and dE/dx (fmmx_energy) against FP64 direct-sum
    values on 1,000 sampled points (the p = 10 truncation: <= 2e-5 and 5e-5),
    then central differences of the energy (h = 1e-4, 64 coordinates, one
    Richardson step; see _pair_difference_fd for how the difference keeps its
    digits): they agree with the FP64 direct-sum gradient to 1e-6
    (normwise), with the FP64 engine's gradient at p = 12 to 1e-6, and with
    the mixed p = 10 gradient to its truncation."""
    n = 1_000_000
    x, q = W.plummer(n, 41)
    xd, qd = _dev(x), _dev(q)
    box = (-10.0, -10.0, -10.0, 20.0)
    f = F.FMM(p=10, lmax=7, precision="mixed", convention="exact", box=box)
    f.set_points(xd)
    E, dq, dx = f.energy(qd)
    idx = torch.from_numpy(np.random.default_rng(43).choice(n, 1000, replace=False)).cuda()
    dv, dg = f.direct(xd[idx], xd, qd)
    gx_ref = qd[idx].double()[:, None] * dg
    e_q, e_x = _rel(dq[idx], dv), _rel(dx[idx], gx_ref)
    print(f"mixed p=10: E {E:.6e}, dE/dq {e_q:.2e}, dE/dx {e_x:.2e}", flush=True)
    assert e_q <= 2e-5 and e_x <= 5e-5
    # central differences at 64 coordinates
    rng = np.random.default_rng(44)
    picks, axes = rng.choice(n, 64, replace=False), rng.integers(0, 3, 64)
    fd = np.array([_pair_difference_fd(f, xd, qd, int(k), int(a), 1e-4) for k, a in zip(picks, axes)])
    pt = torch.from_numpy(picks).cuda()
    _, dgp = f.direct(xd[pt], xd, qd)
    direct = q[picks].astype(np.float64) * dgp.cpu().numpy()[np.arange(64), axes]
    g64 = F.FMM(p=12, lmax=7, precision="fp64", convention="exact", box=box)
    g64.set_points(xd)
    _, _, dx64 = g64.energy(qd)
    fmm64 = dx64[pt].cpu().numpy()[np.arange(64), axes]
    mixed = dx[pt].cpu().numpy()[np.arange(64), axes]
    nd = np.linalg.norm(fd)
    e_dir, e_64, e_mx = (np.linalg.norm(fd - v) / nd for v in (direct, fmm64, mixed))
    print(f"central differences vs direct {e_dir:.2e}, FP64 p=12 {e_64:.2e}, mixed p=10 {e_mx:.2e}", flush=True)
    assert e_dir <= 1e-6 and e_64 <= 1e-6 and e_mx <= 5e-5
    f.close()
    g64.close()


def test_batched_solve_c5():
    """C5: 256 ellipsoids with 20 bumps each, 11.6M panels, the induced-charge
    equation of every surface solved by GMRES to rtol 1e-6 with p = 8 in
    mixed precision, all surfaces in one tree; 4 sampled surfaces against
    their own single solves: energies to 5e-5 (both at p = 8, on different
    trees -- the batch's depth against the single surface's -- so their
    truncation errors differ; measured 1.7e-5)."""
    from old_fmm_bem_amd import bem

    surfaces, charges = W.ellipsoids(256, 0)
    npan = sum(len(fc) for _, fc in surfaces)
    print(f"C5: {len(surfaces)} surfaces, {npan} panels", flush=True)
    assert npan > 10_000_000
    r = bem.batched_solvation(surfaces, charges, p=8, precision="mixed", rtol=1e-6, extra_levels=5)
    print(f"batched: {r['iterations']} iterations, max residual {r['residuals'].max():.1e}", flush=True)
    assert r["residuals"].max() <= 1e-6
    for s in np.random.default_rng(5).choice(len(surfaces), 4, replace=False):
        (v, fc), (xq, qq) = surfaces[s], charges[s]
        one = bem.solvation_energy(v, fc, xq, qq, p=8, precision="mixed", rtol=1e-6)
        print(f"surface {s}: {len(fc)} panels, batched {r['energies'][s]:.6f}, alone {one['energy']:.6f}",
              flush=True)
        assert abs(r["energies"][s] - one["energy"]) <= 5e-5 * abs(one["energy"])
