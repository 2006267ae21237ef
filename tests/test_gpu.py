"""GPU parity of the HIP engine (called through the C ABI of libfmmx.so)
against the reference's own golden vectors and the oracle.

Tolerances:
  - tree, permutations, box lists, interaction lists: bit-exact;
  - FP64 engine vs oracle in the same convention: rel-L2 <= 1e-11 (operators
    and Cartesian harmonics differ from the reference's rotations/Legendre
    recurrences by rounding only);
  - FP64 REF engine vs the compiled reference's float outputs: rel-L2 <= 2e-6
    (the reference accumulates its outputs in float);
  - MIXED engine vs oracle: rel-L2 <= 3e-6 (potential), 1e-5 (gradient);
  - FMM vs FP64 direct sum at p=10: rel-L2 <= 2e-5 (truncation of the
    reference's own algorithm; measured ~1e-5 by the oracle).
"""
import numpy as np
import pytest

from conftest import GOLDEN_CASES, golden, rel_l2, uniform_cube

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no HIP device", allow_module_level=True)

import old_fmm_bem_amd as F  # noqa: E402
from oracle import oracle as O  # noqa: E402


def _engine(d, precision="fp64", keep_lists=True):
    ni, nj, lmax, neq, same, asis, seed = (int(v) for v in d["in_cfg"])
    f = F.FMM(p=int(d["meta"][0]), lmax=lmax, precision=precision, convention="ref", box=tuple(d["in_box"]))
    f.set_points(d["in_xi"], d["in_xj"], ni_normals=d["in_ni"], nj_normals=d["in_nj"], keep_lists=keep_lists)
    return f, lmax, neq


@pytest.mark.parametrize("case", GOLDEN_CASES)
def test_tree_and_lists_bitexact_vs_reference(case):
    d = golden(case)
    f, lmax, neq = _engine(d)
    print(f"case={case} lmax={lmax}", flush=True)
    assert np.array_equal(f.perm("i"), d["nbi"]) and np.array_equal(f.perm("j"), d["nbj"])
    k, a, b = f.boxes("i", lmax)
    assert np.array_equal(k, d["leaf_nfi"]) and np.array_equal(a, d["leaf_ndi0"]) and np.array_equal(b, d["leaf_ndi1"])
    k, a, b = f.boxes("j", lmax)
    assert np.array_equal(k, d["leaf_nfj"]) and np.array_equal(a, d["leaf_ndj0"]) and np.array_equal(b, d["leaf_ndj1"])
    assert np.array_equal(np.concatenate([f.boxes("i", l)[0] for l in range(lmax, 1, -1)]), d["all_nfi"])
    assert np.array_equal(np.concatenate([f.boxes("j", l)[0] for l in range(lmax, 1, -1)]), d["all_nfj"])
    nij, neij = f.interaction_list(lmax, 0)
    assert np.array_equal(nij, d[f"near_nij_{lmax}"]) and np.array_equal(neij, d[f"near_neij_{lmax}"])
    for lev in range(2, lmax + 1):
        nij, neij = f.interaction_list(lev, 1)
        assert np.array_equal(nij, d[f"far_nij_{lev}"]) and np.array_equal(neij, d[f"far_neij_{lev}"]), lev
    f.close()


@pytest.mark.parametrize("case", GOLDEN_CASES)
def test_sort_matches_reference_permutation(case):
    """fmmx_sort: keys and stable sort of one point set as sort/sorti.cxx:12-26
    leaves them (nbi, nbj), including points within ulps of box faces."""
    d = golden(case)
    lmax = int(d["in_cfg"][2])
    f = F.FMM(p=4, lmax=lmax, precision="fp64", convention="ref", box=tuple(d["in_box"]))
    print(f"case={case}", flush=True)
    for x, perm, leaf in ((d["in_xi"], d["nbi"], d["leaf_nfi"]), (d["in_xj"], d["nbj"], d["leaf_nfj"])):
        k, pm = f.sort(x)
        assert np.array_equal(pm, perm)
        assert np.array_equal(np.unique(k), leaf)
        assert np.all(np.diff(k) >= 0)
    # a new domain takes effect at the next call
    f.set_domain(2, tuple(d["in_box"]))
    k, _ = f.sort(d["in_xi"])
    assert k.max() < 64
    f.close()


@pytest.mark.parametrize("case", GOLDEN_CASES)
def test_fp64_ref_matches_reference_outputs(case):
    d = golden(case)
    f, lmax, neq = _engine(d, keep_lists=False)
    print(f"case={case} neq={neq}", flush=True)
    v, g = f.evaluate(d["in_q"], neq=neq)
    assert rel_l2(v, d["fmm_vi"]) < 2e-6
    if neq == 0:
        assert rel_l2(g, np.stack([d["fmm_gxi"], d["fmm_gyi"], d["fmm_gzi"]], 1)) < 2e-6
    if "up_bx" in d:
        bx = f.multipoles().ravel().view(np.float64)
        assert np.abs(bx - d["up_bx"]).max() <= 1e-11 * np.abs(d["up_bx"]).max()
        for lev in range(2, lmax + 1):
            ax = f.locals(lev).ravel().view(np.float64)
            ref = d[f"ax_{lev}"]
            assert np.abs(ax - ref).max() <= 1e-10 * np.abs(ref).max(), lev
    # against the oracle with double outputs: rounding-level agreement
    o = O.OracleFMM(d["in_xi"], d["in_xj"], mp=int(d["meta"][0]), lmax=lmax, neq=neq, conv=O.REF,
                    box=tuple(d["in_box"]), ni_normals=d["in_ni"], nj_normals=d["in_nj"], fp64_out=True)
    ov, og = o.run(d["in_q"])
    assert rel_l2(v, ov) < 1e-11
    if neq == 0:
        assert rel_l2(g, og) < 1e-11
    f.close()


@pytest.mark.parametrize("neq", [0, 1, 2])
@pytest.mark.parametrize("conv", ["ref", "exact"])
@pytest.mark.parametrize("lmax", [3, 4])
def test_fp64_vs_oracle(neq, conv, lmax):
    x, q = uniform_cube(12000, 10 + lmax)
    rng = np.random.default_rng(5)
    nrm = rng.standard_normal((12000, 3)).astype(np.float32)
    print(f"neq={neq} conv={conv} lmax={lmax}", flush=True)
    f = F.FMM(p=10, lmax=lmax, precision="fp64", convention=conv, box=(0, 0, 0, 1))
    f.set_points(x, x, ni_normals=nrm, nj_normals=nrm)
    v, g = f.evaluate(q, neq=neq)
    oc = O.REF if conv == "ref" else O.EXACT
    ov, og = O.fmm(x, x, q, mp=10, lmax=lmax, neq=neq, conv=oc, box=(0, 0, 0, 1), ni_normals=nrm, nj_normals=nrm,
                   fp64_out=True)
    assert rel_l2(v, ov) < 1e-11
    if neq == 0:
        assert rel_l2(g, og) < 1e-11
    f.close()


@pytest.mark.parametrize("neq", [0, 1, 2])
def test_mixed_vs_oracle(neq):
    x, q = uniform_cube(20000, 21)
    rng = np.random.default_rng(6)
    nrm = rng.standard_normal((20000, 3)).astype(np.float32)
    print(f"neq={neq}", flush=True)
    f = F.FMM(p=10, lmax=4, precision="mixed", convention="exact", box=(0, 0, 0, 1))
    f.set_points(x, x, ni_normals=nrm, nj_normals=nrm)
    v, g = f.evaluate(q, neq=neq)
    ov, og = O.fmm(x, x, q, mp=10, lmax=4, neq=neq, conv=O.EXACT, box=(0, 0, 0, 1), ni_normals=nrm, nj_normals=nrm)
    assert rel_l2(v, ov) < 3e-6
    if neq == 0:
        assert rel_l2(g, og) < 1e-5
    f.close()


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_fmm_error_vs_direct_p10(precision):
    x, q = uniform_cube(100000, 31)
    f = F.FMM(p=10, lmax=4, precision=precision, convention="exact", box=(0, 0, 0, 1))
    f.set_timing(True)
    f.set_points(x)
    v, g = f.evaluate(q)
    # every box of the 16^3 leaf grid is occupied: the M2L pair counts of a
    # full cube, 3,096 + 53,352 + 584,136 (levels 2, 3, 4)
    st = f.stats()
    assert st["m2l_pairs"] == 640584 and st["m2l_pairs_leaf"] == 584136, st
    assert 0 < st["ms_m2l_leaf"] <= st["ms_m2l"]
    idx = np.random.default_rng(1).choice(100000, 1000, replace=False)
    dv, dg = f.direct(x[idx], x, q)
    ov, og = O.direct(x[idx], x, q, conv=O.EXACT)
    assert rel_l2(dv, ov) < 1e-12 and rel_l2(dg, og) < 1e-12  # GPU direct == oracle direct
    assert rel_l2(v[idx], dv) < 2e-5
    assert rel_l2(g[idx], dg) < 5e-5
    f.close()


@pytest.mark.parametrize("neq", [0, 1, 2])
@pytest.mark.parametrize("conv", ["ref", "exact"])
def test_gpu_direct_vs_oracle(neq, conv):
    rng = np.random.default_rng(7)
    xi, _ = uniform_cube(700, 41)
    xj, q = uniform_cube(5000, 42)
    ni_n = rng.standard_normal((700, 3)).astype(np.float32)
    nj_n = rng.standard_normal((5000, 3)).astype(np.float32)
    f = F.FMM(p=4, lmax=2, precision="fp64", convention=conv)
    v, g = f.direct(xi, xj, q, neq=neq, ni_normals=ni_n, nj_normals=nj_n)
    ov, og = O.direct(xi, xj, q, neq=neq, conv=O.REF if conv == "ref" else O.EXACT, ni_normals=ni_n, nj_normals=nj_n)
    # REF: the oracle's direct sum stores float outputs like kernel/G_dir.cxx:130-133
    tol = 1e-7 if conv == "ref" else 1e-12
    assert rel_l2(v, ov) < tol
    if neq == 0:
        assert rel_l2(g, og) < tol
    f.close()


@pytest.mark.parametrize("lmax", [2, 5, 6])
def test_depths_and_sparse_trees(lmax):
    """Deep trees with many empty boxes (points on a sphere), and lmax=2."""
    rng = np.random.default_rng(lmax)
    n = 8000
    s = rng.standard_normal((n, 3))
    x = (0.5 + 0.45 * s / np.linalg.norm(s, axis=1, keepdims=True)).astype(np.float32)
    q = rng.random(n).astype(np.float32)
    print(f"lmax={lmax}", flush=True)
    f = F.FMM(p=8, lmax=lmax, precision="fp64", convention="exact", box=(0, 0, 0, 1))
    f.set_points(x, keep_lists=True)
    v, g = f.evaluate(q)
    o = O.OracleFMM(x, x, mp=8, lmax=lmax, conv=O.EXACT, box=(0, 0, 0, 1))
    ov, og = o.run(q)
    assert np.array_equal(f.perm("j"), o.perm()[1])
    for lev in range(2, lmax + 1):
        assert np.array_equal(f.interaction_list(lev, 1)[1], o.interaction_list(lev, 1)[1]), lev
    assert np.array_equal(f.interaction_list(lmax, 0)[1], o.interaction_list(lmax, 0)[1])
    assert rel_l2(v, ov) < 1e-11 and rel_l2(g, og) < 1e-11
    f.close()


def test_edge_cases():
    # single target / single source, disjoint sets, clustered points, empty sides
    rng = np.random.default_rng(9)
    f = F.FMM(p=6, lmax=3, precision="fp64", convention="exact", box=(0, 0, 0, 1))
    xi = np.array([[0.3, 0.4, 0.5]], np.float32)
    xj, q = uniform_cube(500, 50)
    f.set_points(xi, xj)
    v, _ = f.evaluate(q)
    ov, _ = O.fmm(xi, xj, q, mp=6, lmax=3, conv=O.EXACT, box=(0, 0, 0, 1), fp64_out=True)
    assert rel_l2(v, ov) < 1e-11
    f.set_points(xj, xi)
    v, _ = f.evaluate(np.array([2.0], np.float32))
    ov, _ = O.fmm(xj, xi, np.array([2.0], np.float32), mp=6, lmax=3, conv=O.EXACT, box=(0, 0, 0, 1))
    assert rel_l2(v, ov) < 1e-11
    # every point in one leaf box
    xc = (0.01 + 0.02 * rng.random((300, 3))).astype(np.float32)
    qc = rng.random(300).astype(np.float32)
    f.set_points(xc)
    v, _ = f.evaluate(qc)
    dv, _ = O.direct(xc, xc, qc, conv=O.EXACT)
    assert rel_l2(v, dv) < 1e-12  # all near field
    # empty source set and empty target set
    f.set_points(xi, np.zeros((0, 3), np.float32))
    v, g = f.evaluate(np.zeros(0, np.float32))
    assert v.shape == (1,) and v[0] == 0 and np.all(g == 0)
    f.set_points(np.zeros((0, 3), np.float32), xj)
    v, g = f.evaluate(q)
    assert v.shape == (0,)
    f.close()


def test_auto_box_and_level():
    """rd <= 0: root cube from the points (bio/main.cxx:364-380); lmax 0 in
    the REF convention: fmm/nlevel.cxx thresholds."""
    x, q = uniform_cube(120000, 60, lo=-2.0, hi=3.0)
    f = F.FMM(p=6, lmax=0, precision="fp64", convention="ref")
    f.set_points(x)
    assert f.lmax == 3  # 1e5 <= N < 5e5
    assert not f.occupancy()[0].any()  # the cost model did not run
    v, _ = f.evaluate(q)
    box = O.bounding_cube(x)
    ov, _ = O.fmm(x, x, q, mp=6, lmax=3, conv=O.REF, box=box, nthreads=8, fp64_out=True)
    assert rel_l2(v, ov) < 1e-11
    f.close()


def _box_stats_np(x, box, lev):
    """Boxes holding a point and the sum of squared points per box at a
    level, keys as fmm/boxn.cxx forms them (float offsets times the float
    scale widened to double, truncated)."""
    lo = np.asarray(box[:3], np.float32)
    scale = np.float64(np.float32(1 << lev) / np.float32(box[3]))
    c = np.floor((x - lo).astype(np.float32).astype(np.float64) * scale).astype(np.int64).clip(0, (1 << lev) - 1)
    _, cnt = np.unique((c[:, 0] << (2 * lev)) | (c[:, 1] << lev) | c[:, 2], return_counts=True)
    return len(cnt), float((cnt.astype(np.float64) ** 2).sum())


@pytest.mark.parametrize("dist", ["cube", "sphere", "plummer"])
def test_auto_depth_exact(dist):
    """lmax 0 in the EXACT convention: the engine counts the points of every
    level-8 cell on the device (Morton order, each level's counts summed
    from its children's) and applies the cost model of fmmx_auto_lmax to
    the boxes and sums of squared counts.  The statistics equal numpy's
    counts of the same keys at levels 1..8; the depth equals the rule
    applied to them; a uniform 300k cloud at p = 8 gets 4 levels, the
    measured optimum."""
    from old_fmm_bem_amd import workloads as W

    n = 300_000
    x = {"cube": lambda: W.cube(n)[0], "sphere": lambda: W.sphere(n)[0], "plummer": lambda: W.plummer(n)[0]}[dist]()
    f = F.FMM(p=8, lmax=0, precision="mixed", convention="exact")
    f.set_points(x)
    lm, box = f.domain()
    boxes, sumsq = f.occupancy()
    assert boxes[0] == 1 and sumsq[0] == float(n) ** 2
    for lev in range(1, 9):
        nb, sq = _box_stats_np(x, box, lev)
        assert boxes[lev] == nb and sumsq[lev] == sq, lev
    assert lm == F.auto_lmax(boxes, sumsq, 8, "mixed")
    if dist == "cube":
        assert lm == 4
    # the evaluation on the chosen tree is correct
    q = np.random.default_rng(2).random(n).astype(np.float32)
    v, _ = f.evaluate(q)
    idx = np.random.default_rng(3).choice(n, 200, replace=False)
    vd, _ = O.direct(x[idx], x, q, conv=O.EXACT, nthreads=8)
    assert rel_l2(v[idx], vd) < 2e-4  # p = 8 truncation
    f.close()


def test_auto_box_matches_reference():
    """rd <= 0: the engine's root cube equals the one bio/main.cxx:355-368
    computes (tests/golden/ref_box.npz, from the statements compiled with the
    reference's flags), bit for bit, including extents where a float product
    rounds differently."""
    import os

    from conftest import GOLDEN

    d = np.load(os.path.join(GOLDEN, "ref_box.npz"))
    f = F.FMM(p=4, lmax=3, precision="fp64", convention="ref")
    for k in range(len([c for c in d if c.startswith("cloud_")])):
        x = d[f"cloud_{k}"]
        f.set_points(x)
        lm, box = f.domain()
        assert lm == 3
        assert np.array_equal(np.array(box, np.float32), d[f"box_{k}"]), (k, box, d[f"box_{k}"])
    f.close()


def test_invalid_coordinates_and_groups():
    """FMMX_EINVAL for NaN coordinates and for points outside an explicit
    root cube (include/fmmx.h), and for a group level deeper than lmax."""
    x, q = uniform_cube(2000, 71)
    f = F.FMM(p=4, lmax=3, precision="fp64", convention="exact", box=(0, 0, 0, 1))
    for bad in (np.nan, 1.5, -0.25, np.inf):
        xb = x.copy()
        xb[17, 1] = bad
        with pytest.raises(F.FMMError, match="outside the root cube"):
            f.set_points(xb)
        with pytest.raises(F.FMMError):
            f.evaluate(q)  # no tree after a failed set_points
    f.set_points(x)  # recovers
    v, _ = f.evaluate(q)
    assert np.isfinite(v).all()
    g = F.FMM(p=4, lmax=0, precision="fp64", convention="exact")
    xn = x.copy()
    xn[5, 2] = np.nan
    with pytest.raises(F.FMMError, match="non-finite"):
        g.set_points(xn)
    g.close()
    f.set_groups(4)  # deeper than lmax 3
    with pytest.raises(F.FMMError, match="group level"):
        f.set_points(x)
    f.set_groups(2)
    f.set_points(x)
    f.close()
    # an automatic depth is raised to the group level
    a = F.FMM(p=4, lmax=0, precision="fp64", convention="exact")
    a.set_groups(5)
    a.set_points(x)
    assert a.lmax >= 5
    a.close()


def _near_pairs_np(x, box, lev):
    """Particle pairs over the 3x3x3 box neighbourhoods of the leaf boxes."""
    lo = np.asarray(box[:3], np.float32)
    scale = np.float64(np.float32(1 << lev) / np.float32(box[3]))
    c = np.floor((x - lo).astype(np.float32).astype(np.float64) * scale).astype(np.int64).clip(0, (1 << lev) - 1)
    g = 1 << lev
    cnt = np.zeros((g + 2, g + 2, g + 2), np.int64)
    np.add.at(cnt, (c[:, 0] + 1, c[:, 1] + 1, c[:, 2] + 1), 1)
    nb = sum(np.roll(np.roll(np.roll(cnt, dx, 0), dy, 1), dz, 2)
             for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1))
    return int((cnt * nb).sum())


@pytest.mark.parametrize("dist", ["cube", "plummer"])
def test_statistics_after_set_points(dist):
    """The tree statistics are read when asked for: after set_points alone
    the near-pair count equals the pairs of the 3x3x3 neighbourhoods counted
    in numpy, no M2L pairs are reported yet, and after an evaluation the M2L
    pair count is that evaluation's."""
    from old_fmm_bem_amd import workloads as W

    n = 100_000
    x = W.cube(n)[0] if dist == "cube" else W.plummer(n)[0]
    f = F.FMM(p=6, lmax=4, precision="mixed", convention="exact")
    f.set_points(x)
    lm, box = f.domain()
    st = f.stats()
    assert st["p2p_pairs"] == _near_pairs_np(x, box, lm)
    assert st["m2l_pairs"] == 0
    f.evaluate(np.ones(n, np.float32))
    st2 = f.stats()
    assert st2["p2p_pairs"] == st["p2p_pairs"] and st2["m2l_pairs"] > 0
    f.close()


def test_column_inputs_match_rows():
    """set_points with the coordinates as three columns (the engine's own
    layout, no copy) gives the tree and results of the (n, 3) form, bit for
    bit, for host and device arrays."""
    import torch

    x, q = uniform_cube(30000, 77)
    f = F.FMM(p=8, lmax=4, precision="mixed", convention="exact", box=(0, 0, 0, 1))
    f.set_points(x)
    v0, g0 = f.evaluate(q)
    f.set_points(tuple(np.ascontiguousarray(x[:, k]) for k in range(3)))
    v1, g1 = f.evaluate(q)
    assert np.array_equal(v0, v1) and np.array_equal(g0, g1)
    xd = torch.from_numpy(x).cuda()
    f.set_points(tuple(xd[:, k].contiguous() for k in range(3)))
    v2, g2 = f.evaluate(torch.from_numpy(q).cuda())
    assert np.array_equal(v0, v2.cpu().numpy()) and np.array_equal(g0, g2.cpu().numpy())
    with pytest.raises(ValueError):
        f.set_points((xd[:, 0], xd[:, 1]))
    f.close()


def test_device_pointers_match_host():
    x, q = uniform_cube(30000, 70)
    f = F.FMM(p=10, lmax=4, precision="mixed", convention="exact", box=(0, 0, 0, 1))
    f.set_points(x)
    v1, g1 = f.evaluate(q)
    xt = torch.from_numpy(x).cuda()
    qt = torch.from_numpy(q).cuda()
    f.set_points(xt)
    v2, g2 = f.evaluate(qt)
    torch.cuda.synchronize()
    assert np.array_equal(v1, v2.cpu().numpy()) and np.array_equal(g1, g2.cpu().numpy())
    f.close()


def test_deterministic_and_reusable():
    x, q = uniform_cube(40000, 80)
    f = F.FMM(p=10, lmax=4, precision="fp64", convention="exact", box=(0, 0, 0, 1))
    f.set_points(x)
    a, ga = f.evaluate(q)
    b, gb = f.evaluate(q)
    assert np.array_equal(a, b) and np.array_equal(ga, gb)
    # GMRES-style reuse: new charges, same tree; linearity
    q2 = np.random.default_rng(1).random(40000).astype(np.float32)
    c, _ = f.evaluate(q2)
    d, _ = f.evaluate((q + q2).astype(np.float32))
    assert rel_l2(d, a + c) < 1e-7
    f.close()


def test_overlapped_and_serialized_phases_agree():
    """Without per-phase timing the near field runs on a second stream beside
    the far field; with it the phases run in order.  Same bits either way,
    for every kernel."""
    x, q = uniform_cube(50000, 82)
    nrm = np.random.default_rng(3).standard_normal((50000, 3)).astype(np.float32)
    f = F.FMM(p=10, lmax=4, precision="mixed", convention="exact", box=(0, 0, 0, 1), lam=2.0)
    f.set_points(x, x, ni_normals=nrm, nj_normals=nrm)
    for neq in (F.G, F.GNI, F.GNJ, F.YUKAWA):
        print(f"neq={neq}", flush=True)
        f.set_timing(False)
        a = f.evaluate(q, neq=neq)
        f.set_timing(True)
        b = f.evaluate(q, neq=neq)
        assert np.array_equal(a[0], b[0])
        if a[1] is not None:
            assert np.array_equal(a[1], b[1])
    f.close()


@pytest.mark.parametrize("p", [4, 10])
def test_canonical_m2l_operators_match_full_set(p, monkeypatch):
    """The mixed EXACT engine keeps the 56 M2L operators with non-negative
    offsets and signs them per reflection while staging; with
    FMMX_M2L_CANON=0 it uploads all 316.  Both agree to float rounding of the
    operators, for every kernel and at every level."""
    x, q = uniform_cube(60000, 84)
    nrm = np.random.default_rng(4).standard_normal((60000, 3)).astype(np.float32)
    out = []
    for canon in ("0", "1"):
        monkeypatch.setenv("FMMX_M2L_CANON", canon)
        f = F.FMM(p=p, lmax=4, precision="mixed", convention="exact", box=(0, 0, 0, 1))
        f.set_points(x, x, ni_normals=nrm, nj_normals=nrm)
        print(f"p={p} canon={canon}", flush=True)
        out.append([f.evaluate(q, neq=k)[0] for k in (0, 1, 2)] + [f.evaluate(q)[1]])
        f.close()
    for k in range(4):
        assert rel_l2(out[1][k], out[0][k]) < 1e-6, k
    if p == 10:
        idx = np.arange(0, 60000, 200)
        dv, dg = O.direct(x[idx], x, q, neq=0, conv=O.EXACT)
        assert rel_l2(out[1][0][idx], dv) < 2e-5


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_autotune_keeps_results(precision):
    """The autotuner only picks launch shapes and offset splits: results
    after tuning agree with the untuned ones to summation-order rounding."""
    x, q = uniform_cube(200000, 86)
    f = F.FMM(p=10, lmax=5, precision=precision, convention="exact", box=(0, 0, 0, 1))
    f.set_points(x)
    v0, g0 = f.evaluate(q)
    ch = f.autotune(q, reps=1)
    print(f"precision={precision} choices={ch}", flush=True)
    assert set(ch) == {2, 3, 4, 5} and all(1 <= c[3] <= 16 for c in ch.values())
    v1, g1 = f.evaluate(q)
    tol = 1e-12 if precision == "fp64" else 2e-6
    assert rel_l2(v1, v0) < tol and rel_l2(g1, g0) < tol
    f.close()


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_shared_tree_equals_separate_trees(precision):
    """set_points(x) shares one tree between targets and sources; a copy of
    the same points builds two.  Results are bit-identical, then again after
    switching between the two forms on one handle."""
    x, q = uniform_cube(30000, 81)
    nrm = np.random.default_rng(2).standard_normal((30000, 3)).astype(np.float32)
    f = F.FMM(p=8, lmax=4, precision=precision, convention="exact", box=(0, 0, 0, 1))
    out = []
    for xj in (None, x.copy(), None):
        f.set_points(x, xj, ni_normals=nrm, nj_normals=nrm)
        print(f"precision={precision} shared={xj is None}", flush=True)
        out.append([f.evaluate(q, neq=k)[0] for k in (0, 1, 2)] + [f.evaluate(q)[1]])
    for k in range(4):
        assert np.array_equal(out[0][k], out[1][k]) and np.array_equal(out[0][k], out[2][k]), k
    f.close()


def test_error_paths():
    f = F.FMM(p=6, lmax=3, convention="exact", box=(0, 0, 0, 1))
    with pytest.raises(F.FMMError):
        f.evaluate(np.ones(3, np.float32))  # before set_points
    x, q = uniform_cube(100, 90)
    f.set_points(x)
    with pytest.raises(F.FMMError):
        f.evaluate(q, neq=1)  # Gni without target normals
    with pytest.raises(F.FMMError):
        f.evaluate(q, neq=5)
    f.close()
