"""Many independent problems in one FMM (fmmx_set_groups) and the batched
BEM solves built on it (old_fmm_bem_amd.bem.BatchedInducedCharge,
gmres_batched, batched_solvation).  No reference counterpart: the
reference solves one surface per run; each batched result is checked
against the same problem solved alone."""
import numpy as np
import pytest
import torch

from conftest import rel_l2
import old_fmm_bem_amd as F
from old_fmm_bem_amd import bem
from oracle import oracle as O


def test_ellipsoid_surface_geometry():
    v, f = bem.ellipsoid_surface(2000, (1.5, 1.0, 0.6), bumps=[((0, 0, 1), 0.1, 0.3)], seed=1)
    assert abs(len(f) - 2000) <= 4 and f.min() == 1 and f.max() == len(v)
    from old_fmm_bem_amd.formats import panels

    c, area, nrm = panels(v, f)
    # the reference's normal (v2 - v0) x (v1 - v0) points inward for this winding
    assert np.all(np.einsum("ij,ij->i", nrm, c) < 0)
    vol = np.sum(np.einsum("ij,ij->i", c, -nrm) * area) / 3
    assert abs(vol - 4 / 3 * np.pi * 1.5 * 1.0 * 0.6) < 0.05 * vol  # bump adds a little


def _block_system(sizes, seed=0):
    rng = np.random.default_rng(seed)
    blocks = [np.eye(m) * 3 + rng.standard_normal((m, m)) / np.sqrt(m) for m in sizes]
    b = rng.standard_normal(sum(sizes))
    seg = np.repeat(np.arange(len(sizes)), sizes)
    return blocks, b, seg


def _check_gmres_batched(dev):
    # segments shorter than, equal to and longer than the chunk, an empty one
    sizes, chunk = [5, 16, 37, 0, 64, 3], 16
    blocks, b, seg = _block_system(sizes)
    Ab = [torch.from_numpy(a).to(dev) for a in blocks]
    off = np.concatenate([[0], np.cumsum(sizes)])

    def matvec(x):
        return torch.cat([Ab[s] @ x[off[s]:off[s + 1]] for s in range(len(sizes))])

    bt = torch.from_numpy(b).to(dev)
    x, res, it = bem.gmres_batched(matvec, bt, torch.from_numpy(seg).to(dev), len(sizes),
                                   precond=lambda v: v / 3, rtol=1e-10, restart=8, maxiter=200, chunk=chunk)
    x = x.cpu().numpy()
    for s, a in enumerate(blocks):
        if sizes[s]:
            ref = np.linalg.solve(a, b[off[s]:off[s + 1]])
            assert np.linalg.norm(x[off[s]:off[s + 1]] - ref) <= 1e-8 * np.linalg.norm(ref)
    assert float(res.max()) <= 1e-10 and it > 8  # restarted at least once


def test_gmres_batched_block_solves_cpu():
    _check_gmres_batched("cpu")


def test_gmres_batched_rejects_interleaved_segments():
    seg = torch.tensor([0, 1, 0, 1])
    with pytest.raises(ValueError):
        bem.gmres_batched(lambda v: v, torch.ones(4, dtype=torch.float64), seg, 2)


# ---- GPU ----------------------------------------------------------------------
# torch sees the device first (the batched solver keeps its vectors there)
HAVE_GPU = torch.cuda.is_available()


def _cloud(n, centre, size, seed):
    rng = np.random.default_rng(seed)
    return (centre + (rng.random((n, 3)) - 0.5) * size).astype(np.float32)


@pytest.mark.gpu
@pytest.mark.parametrize("G", [1, 2])
def test_groups_isolate_problems(G):
    """Clouds in different cells of the level-G grid: with groups set, each
    cloud's potential is its own direct sum (FMM accuracy); the cross
    interaction, which is of order one here, is absent."""
    cells = 1 << G
    rng = np.random.default_rng(90 + G)
    picks = rng.choice(cells ** 3, 5 if G == 2 else 3, replace=False)
    xs, qs = [], []
    for k, c in enumerate(picks):
        centre = (np.array([c % cells, (c // cells) % cells, c // cells ** 2]) + 0.5) / cells
        xs.append(_cloud(4000, centre, 0.9 / cells, 100 + k))
        qs.append(np.random.default_rng(200 + k).uniform(-1, 1, 4000).astype(np.float32))
    x, q = np.concatenate(xs), np.concatenate(qs)
    f = F.FMM(p=10, lmax=G + 3, precision="fp64", convention="exact", box=(0, 0, 0, 1))
    f.set_groups(G)
    f.set_points(x)
    print(f"G={G} clouds={len(picks)}", flush=True)
    v, g = f.evaluate(q)
    o = 0
    for xc, qc in zip(xs, qs):
        dv, dg = O.direct(xc, xc, qc, conv=O.EXACT)
        sl = slice(o, o + len(xc))
        assert rel_l2(v[sl], dv) < 2e-5 and rel_l2(g[sl], dg) < 5e-5
        full, _ = O.direct(xc, x, q, conv=O.EXACT)
        assert rel_l2(full, dv) > 1e-2  # the other clouds would matter
        o += len(xc)
    f.set_groups(0)
    f.set_points(x)
    v0, _ = f.evaluate(q)
    assert rel_l2(v0[:len(xs[0])], O.direct(xs[0], x, q, conv=O.EXACT)[0]) < 2e-5
    f.close()


@pytest.mark.gpu
def test_batched_solvation_matches_single_solves():
    """Eight ellipsoids of 400 to 6,000 panels, each with its own charges: the
    batched solve (one FMM per GMRES iteration for all of them, per-surface
    Krylov processes) gives each surface the energy of its own solve."""
    rng = np.random.default_rng(95)
    surfaces, charges = [], []
    for s in range(8):
        n = int(np.exp(rng.uniform(np.log(400), np.log(6000))))
        axes = rng.uniform(0.5, 2.0, 3)
        bumps = [(u / np.linalg.norm(u), rng.uniform(0, 0.1), 0.3) for u in rng.standard_normal((4, 3))]
        surfaces.append(bem.ellipsoid_surface(n, axes, bumps, seed=s))
        xq = (rng.uniform(-0.3, 0.3, (3, 3)) * axes).astype(np.float32)
        charges.append((xq, rng.choice([-1.0, 1.0], 3).astype(np.float32)))
    r = bem.batched_solvation(surfaces, charges, p=10, precision="fp64", rtol=1e-9, extra_levels=3)
    print(f"batched: {r['iterations']} iterations, residuals <= {r['residuals'].max():.1e}", flush=True)
    assert r["residuals"].max() <= 1e-9
    for s, ((v, f), (xq, q)) in enumerate(zip(surfaces, charges)):
        one = bem.solvation_energy(v, f, xq, q, p=10, precision="fp64", rtol=1e-9)
        print(f"surface {s}: {len(f)} panels, batched {r['energies'][s]:.6f}, alone {one['energy']:.6f}",
              flush=True)
        assert abs(r["energies"][s] - one["energy"]) < 1e-5 * abs(one["energy"])



@pytest.mark.gpu
def test_gmres_batched_block_solves_gpu():
    """The device path: per-iteration flags read one iteration late."""
    _check_gmres_batched("cuda")
