"""Input and output files of the reference's BEM applications
(old_fmm_bem_amd.formats) and the constant-element solvation pipeline of
bio/main.cxx (old_fmm_bem_amd.bem.solvation_energy).

The input fixtures under tests/golden/bio/ are the reference's own data files
(bio/data/sphere.*, bio/data/tripeptide.*).  The pipeline is pinned by
Kirkwood's closed form for charges inside a dielectric sphere and by a dense
FP64 solve of the same discrete equations built from the oracle's direct
sums."""
import os
import struct

import numpy as np
import pytest
from scipy.special import eval_legendre

from conftest import ROOT, rel_l2
from old_fmm_bem_amd import bem, formats
from oracle import oracle as O

DATA = os.path.join(ROOT, "tests", "golden", "bio")


def kirkwood(xq, q, a, e1, e2, nmax=200):
    """Reaction-field energy (kcal/mol, charges in e, lengths in Angstrom) of
    point charges inside a sphere of radius a: permittivity e1 inside, e2
    outside.  phi_reac(r) = sum_n B_n r^n P_n(cos g) with
    B_n = (n+1)(e1-e2) q r_s^n / (e1 (n e1 + (n+1) e2) a^(2n+1))."""
    xq = np.asarray(xq, float)
    r = np.linalg.norm(xq, axis=1)
    rs = np.where(r > 0, r, 1.0)  # a charge at the centre enters through n = 0 only
    cosg = np.clip((xq @ xq.T) / np.outer(rs, rs), -1, 1)
    rr = np.outer(r, r) / a ** 2
    e = 0.0
    for n in range(nmax):
        c = (n + 1) * (e1 - e2) / (e1 * (n * e1 + (n + 1) * e2)) / a
        e += c * np.sum(np.outer(q, q) * rr ** n * eval_legendre(n, cosg))
    return 0.5 * bem.KCAL * e


def sphere_panels(subdiv, radius):
    """Flat triangles of an icosphere, listed counter-clockwise from outside
    (the winding of bio/data/sphere.face), 1-based."""
    m = bem.icosphere(subdiv, radius)
    V = m.nodes.astype(np.float32)
    F = m.tri[:, :3].copy()
    v = V[F].astype(float)
    out = np.einsum("ij,ij->i", np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]), v.mean(1)) > 0
    F[~out] = F[~out][:, [0, 2, 1]]
    return V, F + 1


# ---- readers -------------------------------------------------------------------
def test_read_reference_sphere():
    vert, face = formats.read_surface(os.path.join(DATA, "sphere"))
    xq, q = formats.read_xyzq(os.path.join(DATA, "sphere.xyzq"))
    assert vert.dtype == np.float32 and vert.shape == (290, 3)
    assert face.shape == (576, 3) and face.min() == 1 and face.max() == 290
    assert xq.shape == (20, 3) and np.all(q == 1)
    assert np.allclose(np.linalg.norm(vert, axis=1), 5.0, atol=2e-3)
    # closed surface: every edge shared by exactly two faces, consistently wound
    e = np.concatenate([face[:, [0, 1]], face[:, [1, 2]], face[:, [2, 0]]])
    assert len({tuple(x) for x in e}) == len(e) and all((b, a) in {tuple(x) for x in e} for a, b in e[:50])


def test_read_tripeptide_and_degenerate_faces(tmp_path):
    vert, face = formats.read_surface(os.path.join(DATA, "tripeptide"))
    raw = formats.read_face(os.path.join(DATA, "tripeptide.face"))
    xq, q = formats.read_xyzq(os.path.join(DATA, "tripeptide.xyzq"))
    assert vert.shape == (1986, 3) and raw.shape == (3968, 3) and xq.shape == (44, 3)
    assert abs(q.sum() + 0.33) < 1e-4  # net charge of the listed partial charges
    # the faces kept are the ones above the area threshold, in file order
    v = vert[raw - 1].astype(float)
    area = 0.5 * np.linalg.norm(np.cross(v[:, 2] - v[:, 0], v[:, 1] - v[:, 0]), axis=1)
    assert np.array_equal(face, raw[area > 1e-5])
    # a planted zero-area face (repeated vertex) is dropped, order kept
    (tmp_path / "t.vert").write_text("0 0 0 1 0 0\n1 0 0\n0 1 0 9 9\n0 0 1\n")
    (tmp_path / "t.face").write_text("1 2 3 1 1\n1 1 4 1 1\n1 3 4\n")
    vt, ft = formats.read_surface(str(tmp_path / "t"))
    assert vt.shape == (4, 3) and ft.tolist() == [[1, 2, 3], [1, 3, 4]]
    (tmp_path / "b.face").write_text("0 1 2\n")
    with pytest.raises(ValueError):
        formats.read_face(str(tmp_path / "b.face"))


def test_six_node_mesh_numbering():
    """Two triangles sharing an edge: element 0 takes nodes 0-5, the shared
    vertices and the shared edge midpoint keep their numbers."""
    vert = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0.5]], np.float32)
    face = np.array([[1, 2, 3], [2, 4, 3]])
    nodes, tri6 = formats.six_node_mesh(vert, face)
    assert tri6.tolist() == [[0, 1, 2, 3, 4, 5], [1, 6, 2, 7, 8, 4]]
    assert np.array_equal(nodes[4], (vert[1] + vert[2]) * np.float32(0.5))
    # the reference sphere: V + E nodes (Euler), every node used
    vert, face = formats.read_surface(os.path.join(DATA, "sphere"))
    nodes, tri6 = formats.six_node_mesh(vert, face)
    assert len(nodes) == 290 + 576 * 3 // 2 and np.array_equal(np.unique(tri6), np.arange(len(nodes)))
    assert np.array_equal(nodes[tri6[:, :3]], vert[face - 1])


def test_panels_match_reference_formulas():
    vert, face = formats.read_surface(os.path.join(DATA, "sphere"))
    c, area, nrm = formats.panels(vert, face)
    assert np.allclose(np.linalg.norm(nrm, axis=1), 1, atol=1e-6)
    assert np.all(np.einsum("ij,ij->i", nrm, c) < 0)  # inward for this winding
    assert abs(area.sum() - 4 * np.pi * 25) < 0.02 * 4 * np.pi * 25


# ---- writers -------------------------------------------------------------------
def test_test_dat_round_trip(tmp_path):
    recs = [(1000, np.arange(100.0), 1.5e-5), (2000, np.ones(100), 2.5e-6)]
    formats.write_test_dat(str(tmp_path / "test.dat"), recs, nprocs=1)
    assert os.path.getsize(tmp_path / "test.dat") == 4 + 2 * (4 + 800 + 8)
    nprocs, back = formats.read_test_dat(str(tmp_path / "test.dat"))
    assert nprocs == 1 and [r[0] for r in back] == [1000, 2000]
    assert np.array_equal(back[0][1], recs[0][1]) and back[1][2] == 2.5e-6


def test_bio_dat_layouts(tmp_path):
    vert, face = formats.read_surface(os.path.join(DATA, "sphere"))
    xq, q = formats.read_xyzq(os.path.join(DATA, "sphere.xyzq"))
    c, area, nrm = formats.panels(vert, face)
    nwn, nwe, nch = len(vert), len(face), len(xq)
    si = np.linspace(0, 1, nwe)
    p0 = str(tmp_path / "bio0.dat")
    formats.write_bio0_dat(p0, vert, face, si, xq, q, c, nrm)
    data = open(p0, "rb").read()
    assert len(data) == 16 + 3 * nwn * 8 + 3 * nwe * 4 + nwe * 8 + 4 * nch * 4 + 6 * nwe * 4
    assert struct.unpack_from("<4i", data) == (1, nwe, nwn, nch)
    rec = np.frombuffer(data[16:16 + 3 * nwn * 8], "<f4").reshape(-1, 2)
    assert np.array_equal(rec[:, 0], vert.reshape(-1)) and np.array_equal(rec[:-1, 1], vert.reshape(-1)[1:])
    off = 16 + 3 * nwn * 8
    assert np.array_equal(np.frombuffer(data[off:off + 12 * nwe], "<i4").reshape(-1, 3), face)
    # the 6-node file
    nodes, tri6 = formats.six_node_mesh(vert, face)
    m = bem.Mesh(nodes, tri6)
    vta, vsa = m.frames()
    assert np.allclose(np.einsum("ij,ij->i", vta, m.normals), 0, atol=1e-12)
    assert np.allclose(np.einsum("ij,ij->i", vsa, m.normals), 0, atol=1e-12)
    p1 = str(tmp_path / "bio.dat")
    formats.write_bio_dat(p1, nodes, tri6, m.normals, vta, vsa, xq, q, np.arange(len(nodes), dtype=float))
    n6 = 6 * nwe
    data = open(p1, "rb").read()
    assert len(data) == 16 + 3 * 4 * len(nodes) + 9 * 8 * len(nodes) + 4 * 4 * nch + n6 * (3 * 8 + 8)
    assert struct.unpack_from("<4i", data) == (nwe, len(nodes), 6, nch)
    assert np.array_equal(np.frombuffer(data[len(data) - 8 * n6:], "<f8"), tri6.reshape(-1).astype(float))


def test_kirkwood_reduces_to_born():
    """One centred charge: the series is Born's (e1 - e2)/(e1 e2 a) term."""
    e = kirkwood(np.zeros((1, 3)), np.array([1.0]), 2.0, 4.0, 20.0)
    assert e == pytest.approx(0.5 * bem.KCAL * (4.0 - 20.0) / (4.0 * 20.0 * 2.0), rel=1e-14)


def _dense_pipeline(vert, face, xq, q, e1, e2):
    """The same discrete equations in FP64 with the oracle's direct sums and a
    dense solve."""
    c, area, nrm = formats.panels(vert, face)
    vi, _ = O.direct(c, xq, q, neq=O.NEQ_GNI, conv=O.EXACT, ni_normals=nrm)
    un = -4 * np.pi * vi / e1
    n = len(c)
    A = np.zeros((n, n))
    for m in range(n):
        w = np.zeros(n, np.float32)
        w[m] = area[m] / e1
        A[:, m], _ = O.direct(c, c, w, neq=O.NEQ_GNI, conv=O.EXACT, ni_normals=nrm)
    A += np.eye(n) * (e1 + e2) / (2 * e1 * (e1 - e2))
    si = np.linalg.solve(A, un)
    vr, _ = O.direct(xq, c, (si * area / e1).astype(np.float32), conv=O.EXACT)
    return 0.5 * bem.KCAL * np.sum(q * vr), si


def test_dense_pipeline_converges_to_kirkwood():
    """The discrete equations of bio/main.cxx solved densely: the energy of an
    off-centre charge approaches Kirkwood's as the panels shrink (320 and
    1,280 panels), and the reference's own sphere (576 panels, radius 5, 20
    unit charges up to 0.3 from the surface) lands within its discretisation
    error."""
    xq, q = np.array([[0.0, 0.0, 2.5]], np.float32), np.ones(1, np.float32)
    ek = kirkwood(xq, q, 5.0, 4.0, 20.0)
    errs = []
    for s in (2, 3):
        vert, face = sphere_panels(s, 5.0)
        errs.append(abs(_dense_pipeline(vert, face, xq, q, 4.0, 20.0)[0] - ek) / abs(ek))
    assert errs[1] < errs[0] / 2 and errs[1] < 0.03, errs
    vert, face = formats.read_surface(os.path.join(DATA, "sphere"))
    xq, q = formats.read_xyzq(os.path.join(DATA, "sphere.xyzq"))
    e, _ = _dense_pipeline(vert, face, xq, q, 4.0, 20.0)
    ek = kirkwood(xq, q, 5.0, 4.0, 20.0)
    assert abs(e - ek) < 0.1 * abs(ek), (e, ek)


# ---- GPU ----------------------------------------------------------------------
@pytest.mark.gpu
def test_solvation_fmm_matches_dense_on_reference_sphere():
    vert, face = formats.read_surface(os.path.join(DATA, "sphere"))
    xq, q = formats.read_xyzq(os.path.join(DATA, "sphere.xyzq"))
    e_dense, si_dense = _dense_pipeline(vert, face, xq, q, 4.0, 20.0)
    r = bem.solvation_energy(vert, face, xq, q, 4.0, 20.0, p=12, precision="fp64", rtol=1e-10)
    print(f"fmm {r['energy']:.6f} dense {e_dense:.6f} iterations {r['iterations']}", flush=True)
    # the FMM's truncation error of the normal-derivative kernel (measured
    # 2.6e-5 at p = 10) bounds the agreement of the panel charges
    assert rel_l2(r["si"], si_dense) < 2e-5
    assert abs(r["energy"] - e_dense) < 2e-6 * abs(e_dense)


@pytest.mark.gpu
def test_solvation_converges_to_kirkwood():
    """Off-centre charges in icosphere panel meshes of radius 5 (1,280 to
    20,480 panels): the error against Kirkwood's series falls with the panel
    size and is below 1% on the finest mesh."""
    rng = np.random.default_rng(11)
    xq = (rng.random((30, 3)) - 0.5) * 5
    xq = xq[np.linalg.norm(xq, axis=1) < 3.5].astype(np.float32)
    q = rng.choice([-1.0, 1.0], len(xq)).astype(np.float32)
    ek = kirkwood(xq, q, 5.0, 4.0, 20.0)
    errs = []
    for s in (3, 4, 5):
        vert, face = sphere_panels(s, 5.0)
        r = bem.solvation_energy(vert, face, xq, q, 4.0, 20.0, p=10, precision="mixed")
        errs.append(abs(r["energy"] - ek) / abs(ek))
        print(f"subdiv {s}: {len(face)} panels, energy {r['energy']:.5f} (Kirkwood {ek:.5f}), "
              f"{r['iterations']} iterations", flush=True)
    assert errs[2] < errs[1] < errs[0] and errs[2] < 0.01, errs


@pytest.mark.gpu
def test_solvation_tripeptide_mixed_vs_fp64():
    """The reference's tripeptide surface and charges: the mixed-precision
    pipeline agrees with the FP64 one to the GMRES tolerance."""
    vert, face = formats.read_surface(os.path.join(DATA, "tripeptide"))
    xq, q = formats.read_xyzq(os.path.join(DATA, "tripeptide.xyzq"))
    a = bem.solvation_energy(vert, face, xq, q, p=10, precision="fp64", rtol=1e-8)
    b = bem.solvation_energy(vert, face, xq, q, p=10, precision="mixed", rtol=1e-8)
    print(f"tripeptide: fp64 {a['energy']:.6f} mixed {b['energy']:.6f} kcal/mol", flush=True)
    assert a["energy"] < 0
    assert abs(a["energy"] - b["energy"]) < 1e-4 * abs(a["energy"])


@pytest.mark.gpu
def test_six_node_pipeline_centred_charge():
    """biobem/main.cxx on quadratic elements: for a charge at the centre of
    a sphere the right-hand side is constant, the singularity-subtracted
    double layer annihilates constants, and the energy is
    0.5 * 332.112 * q^2 (e1 - e2) / (e1 (e1 + e2)) * area / (4 pi a^2) / a:
    e2 / (e1 + e2) of Born's energy, which is what these equations give
    (the constant-element equations of bio/main.cxx give Born's)."""
    a, e1, e2 = 5.0, 4.0, 20.0
    m = bem.icosphere(3, a)
    xq, q = np.zeros((1, 3), np.float32), np.array([1.5], np.float32)
    r = bem.solvation_energy_6node(m, xq, q, e1, e2, p=10, precision="fp64", rtol=1e-10)
    area = m.qw.sum()
    want = 0.5 * bem.KCAL * 1.5 ** 2 * (e1 - e2) / (e1 * (e1 + e2)) * area / (4 * np.pi * a ** 2) / a
    print(f"6-node energy {r['energy']:.6f}, closed form {want:.6f}", flush=True)
    # the nodal normals of the quadratic mesh are radial to ~1e-4
    assert abs(r["energy"] - want) < 2e-3 * abs(want)
    assert np.std(r["si"]) < 2e-3 * abs(np.mean(r["si"]))
