"""Surface, charge and result files of the reference's BEM applications.

Inputs (text, one record per line, extra columns ignored):
  <name>.vert   x y z [nx ny nz ...]       vertices (biobem/geometry.cxx:25-37)
  <name>.face   i j k [...]                triangles, 1-based vertex numbers
                                           (biobem/geometry.cxx:38-64)
  <name>.xyzq   x y z q                    point charges (biobem/geometry.cxx:65-78)
Values are parsed as double and kept in single precision, as the reference
stores them.  Faces whose area is at most 1e-5 are dropped on reading
(biobem/geometry.cxx:46-63, bio/main.cxx:164-182).

Outputs (binary, native little-endian):
  bio0.dat   the constant-element application (bio/main.cxx:511-535)
  bio.dat    the 6-node application (biobem/main.cxx:203-238)
  time.dat   timing record (bio/main.cxx:438-443)
  test.dat   the FMM test driver (test/test.cxx:69-70,148-160)
"""
from __future__ import annotations

import numpy as np


def _columns(path, ncol, dtype):
    rows = []
    with open(path) as f:
        for line in f:
            parts = line.split()
            if not parts:
                continue
            if len(parts) < ncol:
                raise ValueError(f"{path}: line {len(rows) + 1} has {len(parts)} columns, {ncol} needed")
            rows.append([dtype(v) for v in parts[:ncol]])
    return np.array(rows, dtype=float if dtype is float else np.int64).reshape(-1, ncol)


def read_vert(path):
    """(nwn, 3) float32 vertex coordinates."""
    return _columns(path, 3, float).astype(np.float32)


def read_face(path):
    """(n, 3) int64 triangles, 1-based as in the file."""
    f = _columns(path, 3, int)
    if f.size and f.min() < 1:
        raise ValueError(f"{path}: vertex numbers are 1-based")
    return f


def read_xyzq(path):
    """(positions (nch, 3) float32, charges (nch,) float32)."""
    a = _columns(path, 4, float).astype(np.float32)
    return a[:, :3].copy(), a[:, 3].copy()


def drop_degenerate_faces(vert, face, tol=1e-5):
    """Faces with area above tol, in file order (biobem/geometry.cxx:46-63:
    the area is half the cross product of two float edge vectors, formed in
    double)."""
    face = np.asarray(face, np.int64)
    if face.size and face.max() > len(vert):
        raise ValueError("face refers to a vertex beyond the vertex list")
    v = np.asarray(vert, np.float32)[face - 1].astype(np.float64)
    b, d = v[:, 2] - v[:, 0], v[:, 1] - v[:, 0]
    area = 0.5 * np.linalg.norm(np.cross(b, d), axis=1)
    return face[area > tol]


def read_surface(base):
    """<base>.vert and <base>.face -> (vertices, faces without degenerate
    ones, 1-based)."""
    vert = read_vert(base + ".vert")
    return vert, drop_degenerate_faces(vert, read_face(base + ".face"))


def six_node_mesh(vert, face, eps=1e-6):
    """Quadratic elements from flat triangles (biobem/geometry.cxx:80-135):
    nodes 3, 4, 5 at the float midpoints of edges 0-1, 1-2, 2-0; global nodes
    numbered in order of first appearance, element 0's six first, a node
    matching an earlier one within eps in every coordinate taking the number
    of the last such node.  Returns (nodes (nwn, 3) float32, tri6 (nwe, 6))."""
    v = np.asarray(vert, np.float32)[np.asarray(face, np.int64) - 1]  # (nwe, 3, 3)
    half = np.float32(0.5)
    x6 = np.concatenate([v, half * (v[:, [0, 1, 2]] + v[:, [1, 2, 0]])], axis=1)  # float sums, as the reference
    nwe = len(x6)
    tri6 = np.zeros((nwe, 6), np.int64)
    if nwe == 0:
        return np.zeros((0, 3), np.float32), tri6
    pts = list(x6[0])
    tri6[0] = np.arange(6)
    # candidate matches by a grid of cell eps; the reference's test is
    # |a - b| <= eps per coordinate, so neighbouring cells are searched too
    cell = {}

    def key(p):
        return tuple(np.floor(p.astype(np.float64) / eps).astype(np.int64))

    for k, p in enumerate(pts):
        cell.setdefault(key(p), []).append(k)
    offs = [(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)]
    for i in range(1, nwe):
        for j in range(6):
            p = x6[i, j]
            kp = key(p)
            best = -1
            for o in offs:
                for k in cell.get((kp[0] + o[0], kp[1] + o[1], kp[2] + o[2]), ()):
                    if k > best and np.all(np.abs(pts[k].astype(np.float64) - p.astype(np.float64)) <= eps):
                        best = k
            if best < 0:
                best = len(pts)
                pts.append(p)
                cell.setdefault(kp, []).append(best)
            tri6[i, j] = best
    return np.array(pts, np.float32), tri6


def panels(vert, face):
    """Constant elements of bio/main.cxx:257-281: centroids (float sums of
    vertex/3), areas, and unit normals (v2 - v0) x (v1 - v0) / (2 area),
    which point inward for faces listed counter-clockwise from outside."""
    v = np.asarray(vert, np.float32)[np.asarray(face, np.int64) - 1]
    third = np.float32(3)
    c = v[:, 0] / third + v[:, 1] / third + v[:, 2] / third
    b = (v[:, 2] - v[:, 0]).astype(np.float32)
    d = (v[:, 1] - v[:, 0]).astype(np.float32)
    n = np.cross(b, d).astype(np.float32)
    area = (0.5 * np.sqrt(np.sum(n.astype(np.float64) ** 2, axis=1))).astype(np.float32)
    return c.astype(np.float32), area, (n / area[:, None] / np.float32(2)).astype(np.float32)


# ---- binary outputs -----------------------------------------------------------
def _i32(f, *v):
    f.write(np.asarray(v, "<i4").tobytes())


def write_bio0_dat(path, vert, face, si, xq, vq_energy, centroids, gvec, nprocs=1):
    """bio/main.cxx:511-535: nprocs, nwe, nwn, nch (int32); 3 nwn 8-byte
    records from the float vertex array (record l holds floats l and l+1, as
    the reference writes `&vert[l]` with the size of a double); faces (3 nwe
    int32, 1-based); si (nwe doubles); charge x, y, z and per-charge energies
    (float32 each); centroid x, y, z and the target-gradient buffer x, y, z
    (float32 each; the reference's buffer then holds the reaction-field
    gradient at the nch charges followed by the panel normals)."""
    vert, face = np.asarray(vert, np.float32), np.asarray(face)
    flat = np.concatenate([vert.reshape(-1), np.zeros(1, np.float32)]).astype("<f4")
    with open(path, "wb") as f:
        _i32(f, nprocs, len(face), len(vert), len(xq))
        f.write(np.stack([flat[:-1], flat[1:]], 1).tobytes())
        f.write(np.asarray(face, "<i4").reshape(-1).tobytes())
        f.write(np.asarray(si, "<f8").tobytes())
        for k in range(3):
            f.write(np.asarray(xq[:, k], "<f4").tobytes())
        f.write(np.asarray(vq_energy, "<f4").tobytes())
        for a in (centroids, gvec):
            for k in range(3):
                f.write(np.asarray(a[:, k], "<f4").tobytes())


def _float_pairs(a, k):
    """What `write(&a[i][k], sizeof(double))` emits for rows of 3 floats: the
    8 bytes of a[i][k] and the float after it (the next row's first value
    for k = 2; zero after the last row)."""
    flat = np.concatenate([np.asarray(a, "<f4").reshape(-1), np.zeros(1, "<f4")])
    idx = 3 * np.arange(len(a)) + k
    return np.stack([flat[idx], flat[idx + 1]], 1).reshape(-1).tobytes()


def write_bio_dat(path, nodes, tri6, vna, vta, vsa, xq, vq_energy, si):
    """biobem/main.cxx:203-238: nwe, nwn, nen=6, nch (int32); node x, y, z
    (float32 each); nodal normal, tangent and binormal components written as
    8-byte records from float storage (two floats each, as the reference
    writes them); charge x, y, z and per-charge energies (float32); then per
    element node (6 nwe of them) x, y, z as 8-byte records of the float node
    coordinates and si (double)."""
    nodes, tri6 = np.asarray(nodes, np.float32), np.asarray(tri6, np.int64)
    with open(path, "wb") as f:
        _i32(f, len(tri6), len(nodes), 6, len(xq))
        for k in range(3):
            f.write(np.asarray(nodes[:, k], "<f4").tobytes())
        for a in (vna, vta, vsa):
            for k in range(3):
                f.write(_float_pairs(a, k))
        for k in range(3):
            f.write(np.asarray(xq[:, k], "<f4").tobytes())
        f.write(np.asarray(vq_energy, "<f4").tobytes())
        idx = tri6.reshape(-1)
        pairs = [_float_pairs(nodes, k) for k in range(3)]
        for k in range(3):
            allp = np.frombuffer(pairs[k], "<f4").reshape(-1, 2)
            f.write(allp[idx].tobytes())
        f.write(np.asarray(si, "<f8")[idx].tobytes())


def write_time_dat(path, n, tfmm):
    """n (int32) and the 100 doubles of the timing array."""
    t = np.zeros(100)
    t[:min(100, len(tfmm))] = np.asarray(tfmm, float)[:100]
    with open(path, "wb") as f:
        _i32(f, n)
        f.write(t.astype("<f8").tobytes())


def write_test_dat(path, records, nprocs=1):
    """test/test.cxx: nprocs (int32), then per run np (int32), the 100 timing
    doubles and the rel-L2 error (double)."""
    with open(path, "wb") as f:
        _i32(f, nprocs)
        for n, tfmm, err in records:
            _i32(f, n)
            t = np.zeros(100)
            t[:min(100, len(tfmm))] = np.asarray(tfmm, float)[:100]
            f.write(t.astype("<f8").tobytes())
            f.write(np.float64(err).astype("<f8").tobytes())


def read_test_dat(path):
    """Inverse of write_test_dat: (nprocs, [(np, tfmm, err), ...])."""
    data = open(path, "rb").read()
    nprocs = int(np.frombuffer(data[:4], "<i4")[0])
    recs, p = [], 4
    while p < len(data):
        n = int(np.frombuffer(data[p:p + 4], "<i4")[0])
        t = np.frombuffer(data[p + 4:p + 804], "<f8").copy()
        e = float(np.frombuffer(data[p + 804:p + 812], "<f8")[0])
        recs.append((n, t, e))
        p += 812
    return nprocs, recs
