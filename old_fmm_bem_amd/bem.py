"""Boundary-element layer over the FMM matvec: the collocation operator of
biobem/fmm_matmult.h:9-196 with the geometry of biobem/bem3d.cxx and an
in-repo restarted GMRES (the reference drives PETSc KSP, bio/main.cxx:391-421).

Surfaces are meshed by 6-node (quadratic) triangles: nodes 0, 1, 2 are the
vertices, 3, 4, 5 sit on the edges 0-1, 1-2, 2-0 (biobem/bem3d.cxx:359-377).
Quadrature points carry the interpolated position, the interpolated nodal
normal and the weight hsg = |dx/dxi x dx/deta| * w (biobem/bem3d.cxx:381-394);
the element integral of f is 0.5 * sum_j hsg_j f_j.

The operator applied to nodal values u (biobem/fmm_matmult.h):
    y_l = (1/e1) [ sum_g 0.5 hsg_g (dG/dn_g)(x_l, y_g) u(y_g)
                 + u_l sum_g 0.5 hsg_g (-(dG/dn_g)(x_l, y_g))
                 + (e1 + e2)/(e1 - e2) u_l ]
with G = 1/(4 pi r) and u(y_g) interpolated by the shape functions: the
double layer with its singularity subtracted.  Both sums are one FMM matvec
with the Gnj kernel (targets: nodes, sources: quadrature points); the second
does not depend on u and is computed once.
"""
from __future__ import annotations

import numpy as np

from . import FMM, G, GNI, GNJ

# ---- quadrature on the reference triangle (biobem/bem3d.cxx:9-245) ---------
# Published symmetric rules (Strang-Fix / Dunavant): points (xi, eta) and
# weights normalised to sum 1.
_RULES = {
    1: ([(1 / 3, 1 / 3)], [1.0]),
    3: ([(1 / 6, 1 / 6), (2 / 3, 1 / 6), (1 / 6, 2 / 3)], [1 / 3] * 3),
    4: ([(1 / 3, 1 / 3), (0.2, 0.2), (0.6, 0.2), (0.2, 0.6)], [-27 / 48, 25 / 48, 25 / 48, 25 / 48]),
}


def _six(al, be, ga, de, o1, o2):
    pts = [(de, de), (al, de), (de, al), (be, be), (ga, be), (be, ga)]
    return pts, [o1] * 3 + [o2] * 3


_RULES[6] = _six(0.816847572980459, 0.445948490915965, 0.108103018168070, 0.091576213509771,
                 0.109951743655322, 0.223381589678011)
# 7 points: 0.797426985353087 = 1 - 2 * 0.101286507323456 (the reference has
# 0.797426958353087 at biobem/bem3d.cxx:73, two digits swapped)
_p7, _w7 = _six(0.797426985353087, 0.470142064105115, 0.059715871789770, 0.101286507323456,
                0.125939180544827, 0.132394152788506)
_RULES[7] = (_p7 + [(1 / 3, 1 / 3)], _w7 + [0.225])


def gauss_trgl(ngl=3):
    """(xi, eta, w) of the ngl-point rule; ngl in 1, 3, 4, 6, 7
    (include/constants.h:33 uses 3)."""
    if ngl not in _RULES:
        raise ValueError(f"ngl must be one of {sorted(_RULES)}")
    pts, w = _RULES[ngl]
    pts = np.asarray(pts, float)
    return pts[:, 0].copy(), pts[:, 1].copy(), np.asarray(w, float)


# ---- 6-node triangle (biobem/bem3d.cxx:248-323) ------------------------------
def interp_p(X, al, be, ga, xi, eta):
    """Shape functions, tangents and unit normal of 6-node triangles.
    X: (ne, 6, 3) node coordinates; al, be, ga: (ne,) edge-node parameters
    of nodes 3, 5, 4; xi, eta: scalars.  Returns ph (ne, 6), dxdxi, dxdeta,
    unit normal (ne, 3 each) and the surface metric hsd (ne,)."""
    alc, bec, gac = 1 - al, 1 - be, 1 - ga
    alalc, bebec, gagac = al * alc, be * bec, ga * gac
    ph2 = xi * (xi - al + eta * (al - ga) / gac) / alc
    ph3 = eta * (eta - be + xi * (be + ga - 1) / ga) / bec
    ph4 = xi * (1 - xi - eta) / alalc
    ph5 = xi * eta / gagac
    ph6 = eta * (1 - xi - eta) / bebec
    ph1 = 1 - ph2 - ph3 - ph4 - ph5 - ph6
    d2 = (2 * xi - al + eta * (al - ga) / gac) / alc
    d3 = eta * (be + ga - 1) / (ga * bec)
    d4 = (1 - 2 * xi - eta) / alalc
    d5 = eta / gagac
    d6 = -eta / bebec
    d1 = -d2 - d3 - d4 - d5 - d6
    e2 = xi * (al - ga) / (alc * gac)
    e3 = (2 * eta - be + xi * (be + ga - 1) / ga) / bec
    e4 = -xi / alalc
    e5 = xi / gagac
    e6 = (1 - xi - 2 * eta) / bebec
    e1 = -e2 - e3 - e4 - e5 - e6
    ph = np.stack([ph1, ph2, ph3, ph4, ph5, ph6], 1)
    dxi = np.einsum("ek,ekc->ec", np.stack([d1, d2, d3, d4, d5, d6], 1), X)
    det = np.einsum("ek,ekc->ec", np.stack([e1, e2, e3, e4, e5, e6], 1), X)
    vn = np.cross(dxi, det)
    hsd = np.linalg.norm(vn, axis=1)
    return ph, dxi, det, vn / hsd[:, None], hsd


class Mesh:
    """Nodes (nwn, 3) and 6-node triangles (nwe, 6) with the per-element
    quadrature data of biobem/bem3d.cxx:325-542 (elm_geom)."""

    def __init__(self, nodes, tri6, ngl=3):
        self.nodes = np.asarray(nodes, float)
        self.tri = np.asarray(tri6, np.int64)
        self.ngl = ngl
        X = self.nodes[self.tri]  # (ne, 6, 3)
        dist = lambda a, b: np.linalg.norm(X[:, a] - X[:, b], axis=1)  # noqa: E731
        # edge-node positions along their edges (biobem/bem3d.cxx:360-365)
        al = 1 / (1 + dist(3, 1) / dist(3, 0))
        be = 1 / (1 + dist(5, 2) / dist(5, 0))
        ga = 1 / (1 + dist(4, 1) / dist(4, 2))
        xq, eq, wq = gauss_trgl(ngl)
        ne = len(self.tri)
        self.ph = np.zeros((ne, ngl, 6))
        self.hsg = np.zeros((ne, ngl))
        for j in range(ngl):
            ph, _, _, _, hsd = interp_p(X, al, be, ga, xq[j], eq[j])
            self.ph[:, j] = ph
            self.hsg[:, j] = hsd * wq[j]
        self.area = 0.5 * self.hsg.sum(1)
        # nodal normals: average of the element normals at the node
        # (biobem/bem3d.cxx:396-407, 503-512)
        vna = np.zeros_like(self.nodes)
        cnt = np.zeros(len(self.nodes))
        zero, one = np.zeros(ne), np.ones(ne)
        node_xi_eta = [(zero, zero), (one, zero), (zero, one), (al, zero), (ga, 1 - ga), (zero, be)]
        for i, (xi, eta) in enumerate(node_xi_eta):
            _, _, _, vn, _ = interp_p(X, al, be, ga, xi, eta)
            np.add.at(vna, self.tri[:, i], vn)
            np.add.at(cnt, self.tri[:, i], 1)
        vna /= cnt[:, None]
        self.normals = vna / np.linalg.norm(vna, axis=1)[:, None]
        # quadrature points: interpolated positions and (unnormalised) normals
        # (biobem/fmm_matmult.h:80-98)
        self.qx = np.einsum("egk,ekc->egc", self.ph, X).reshape(-1, 3)
        self.qn = np.einsum("egk,ekc->egc", self.ph, self.normals[self.tri]).reshape(-1, 3)
        self.qw = 0.5 * self.hsg.reshape(-1)

    @property
    def nwn(self):
        return len(self.nodes)

    @property
    def nwe(self):
        return len(self.tri)

    def frames(self):
        """Nodal tangent and binormal (vta, vsa) of biobem/bem3d.cxx:528-541:
        vsa = (0, n_z, -n_y)/|.| (or (0, 1, 0) when n is within 1e-3 of the
        x axis), vta = vsa x n."""
        n = self.normals
        s = np.sqrt(n[:, 1] ** 2 + n[:, 2] ** 2)
        ok = s > 1e-3
        vsa = np.zeros_like(n)
        vsa[:, 1] = np.where(ok, n[:, 2] / np.where(ok, s, 1), 1.0)
        vsa[:, 2] = np.where(ok, -n[:, 1] / np.where(ok, s, 1), 0.0)
        return np.cross(vsa, n), vsa

    def interpolate(self, u):
        """Nodal values -> values at the quadrature points (ne*ngl,)."""
        return np.einsum("egk,ek->eg", self.ph, np.asarray(u)[self.tri]).reshape(-1)


def icosphere(subdiv, radius=1.0, ngl=3):
    """Sphere meshed by 20*4^subdiv 6-node triangles: an icosahedron refined
    subdiv times, edge nodes at the projected edge midpoints."""
    t = (1 + 5 ** 0.5) / 2
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    V = [np.array(x, float) / np.linalg.norm(x) for x in v]
    F = list(f)

    def midpoints(V, F):
        cache = {}

        def mid(a, b):
            k = (min(a, b), max(a, b))
            if k not in cache:
                m = V[a] + V[b]
                V.append(m / np.linalg.norm(m))
                cache[k] = len(V) - 1
            return cache[k]

        return mid

    for _ in range(subdiv):
        mid = midpoints(V, F)
        G = []
        for a, b, c in F:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            G += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        F = G
    mid = midpoints(V, F)
    tri6 = np.array([(a, b, c, mid(a, b), mid(b, c), mid(c, a)) for a, b, c in F], np.int64)
    return Mesh(np.array(V) * radius, tri6, ngl)


class Collocation:
    """The collocation operator of biobem/fmm_matmult.h:9-196 on a mesh, with
    inner and outer dielectric constants e1, e2, applied through the FMM."""

    def __init__(self, mesh: Mesh, e1=1.0, e2=80.0, p=10, precision="mixed", lmax=0):
        self.mesh, self.e1, self.e2 = mesh, float(e1), float(e2)
        self.f = FMM(p=p, lmax=lmax, precision=precision, convention="exact")
        self._xi = np.ascontiguousarray(mesh.nodes, np.float32)
        self._xj = np.ascontiguousarray(mesh.qx, np.float32)
        self._nj = np.ascontiguousarray(mesh.qn, np.float32)
        self.f.set_points(self._xi, self._xj, nj_normals=self._nj)
        # the u-independent sum: sum_g 0.5 hsg (-(dG/dn_g)) at each node
        k1, _ = self.f.evaluate((-mesh.qw).astype(np.float32), neq=GNJ)
        self.k1 = k1
        self.diag = (self.e1 + self.e2) / (self.e1 - self.e2)

    def matvec(self, u):
        u = np.asarray(u, np.float64)
        s = (self.mesh.qw * self.mesh.interpolate(u)).astype(np.float32)
        v, _ = self.f.evaluate(s, neq=GNJ)
        return (v + (self.k1 + self.diag) * u) / self.e1

    def precondition(self, x):
        """biobem/fmm_matmult.h:198-214: scaling by 2 e1 (e1 - e2)/(e1 + e2)."""
        return np.asarray(x) / (self.e1 + self.e2) * (2 * self.e1 * (self.e1 - self.e2))

    def close(self):
        self.f.close()


def gmres(matvec, b, x0=None, precond=None, rtol=1e-6, restart=50, maxiter=500):
    """Right-preconditioned restarted GMRES, FP64 arithmetic, modified
    Gram-Schmidt with one reorthogonalisation pass.  Returns (x, residual
    norms relative to |b|, iterations)."""
    b = np.asarray(b, np.float64)
    M = precond or (lambda v: v)
    x = np.zeros_like(b) if x0 is None else np.array(x0, np.float64)
    nb = np.linalg.norm(b)
    if nb == 0:
        return x, [0.0], 0
    hist, it = [], 0
    r = b - matvec(x)
    beta = np.linalg.norm(r)
    hist.append(beta / nb)
    while hist[-1] > rtol and it < maxiter:
        m = min(restart, maxiter - it)
        V = np.zeros((m + 1, len(b)))
        H = np.zeros((m + 1, m))
        V[0] = r / beta
        g = np.zeros(m + 1)
        g[0] = beta
        cs, sn = np.zeros(m), np.zeros(m)
        k = 0
        for k in range(m):
            w = matvec(M(V[k]))
            for _ in range(2):  # MGS, then once more
                for i in range(k + 1):
                    h = V[i] @ w
                    H[i, k] += h
                    w = w - h * V[i]
            H[k + 1, k] = np.linalg.norm(w)
            if H[k + 1, k] > 0:
                V[k + 1] = w / H[k + 1, k]
            for i in range(k):
                t = cs[i] * H[i, k] + sn[i] * H[i + 1, k]
                H[i + 1, k] = -sn[i] * H[i, k] + cs[i] * H[i + 1, k]
                H[i, k] = t
            d = np.hypot(H[k, k], H[k + 1, k])
            cs[k], sn[k] = H[k, k] / d, H[k + 1, k] / d
            H[k, k] = d
            H[k + 1, k] = 0
            g[k + 1] = -sn[k] * g[k]
            g[k] = cs[k] * g[k]
            it += 1
            hist.append(abs(g[k + 1]) / nb)
            if hist[-1] <= rtol:
                break
        kk = k + 1
        y = np.linalg.solve(np.triu(H[:kk, :kk]), g[:kk])
        x = x + M(V[:kk].T @ y)
        r = b - matvec(x)
        beta = np.linalg.norm(r)
        hist[-1] = beta / nb  # the true residual
        if beta == 0:
            break
    return x, hist, it


def solve_refined(op_low, op_high, b, rtol=1e-10, inner_rtol=1e-5, restart=50, maxiter=400, sweeps=10):
    """Iterative refinement: residuals with the FP64 operator op_high, each
    correction solved by GMRES on the mixed-precision op_low."""
    x = np.zeros_like(np.asarray(b, np.float64))
    nb = np.linalg.norm(b)
    hist = []
    for _ in range(sweeps):
        r = b - op_high.matvec(x)
        hist.append(np.linalg.norm(r) / nb)
        if hist[-1] <= rtol:
            break
        d, _, _ = gmres(op_low.matvec, r, precond=op_low.precondition, rtol=inner_rtol, restart=restart,
                        maxiter=maxiter)
        x = x + d
    return x, hist


# ---- the constant-element application (bio/main.cxx) -------------------------
# Energy unit: kcal/mol for charges in e and lengths in Angstrom, the constant
# the reference multiplies by (bio/main.cxx:502, biobem/main.cxx:198).
KCAL = 332.112


class InducedCharge:
    """The operator of bio/main.cxx:27-53 on flat panels (centroids c, unit
    normals nrm as bio/main.cxx:257-281 forms them, areas): for panel values x,
        y_l = sum_m Gni(c_l, c_m) x_m area_m / e1 + (e1 + e2)/(2 e1 (e1 - e2)) x_l,
    one FMM matvec with the target-normal kernel (neq 1) per application, and
    the shell preconditioner of bio/main.cxx:55-71."""

    def __init__(self, centroids, normals, area, e1=4.0, e2=20.0, p=10, precision="mixed", convention="exact",
                 lmax=0):
        self.c = np.ascontiguousarray(centroids, np.float32)
        self.nrm = np.ascontiguousarray(normals, np.float32)
        self.area = np.asarray(area, np.float64)
        self.e1, self.e2 = float(e1), float(e2)
        self.f = FMM(p=p, lmax=lmax, precision=precision, convention=convention)
        self.f.set_points(self.c, ni_normals=self.nrm)
        self.diag = (self.e1 + self.e2) / (2 * self.e1 * (self.e1 - self.e2))

    def matvec(self, x):
        x = np.asarray(x, np.float64)
        v, _ = self.f.evaluate((x * self.area / self.e1).astype(np.float32), neq=GNI)
        return v + self.diag * x

    def precondition(self, x):
        return np.asarray(x) / (self.e1 + self.e2) * (2 * self.e1 * (self.e1 - self.e2))

    def close(self):
        self.f.close()


def solvation_energy(vert, face, xq, q, e1=4.0, e2=20.0, p=10, precision="mixed", rtol=1e-5, restart=30,
                     maxiter=1000):
    """Electrostatic solvation energy of point charges inside a triangulated
    surface, the pipeline of bio/main.cxx:256-503 with the FMM of this
    package in place of fmm():
      1. panels: centroids, areas, normals (bio/main.cxx:257-281);
      2. right-hand side un = -4 pi Gni(charges -> centroids) / e1
         (bio/main.cxx:282-345);
      3. GMRES(restart 30, rtol 1e-5: the PETSc KSP defaults) on
         InducedCharge with its shell preconditioner, zero initial guess
         (bio/main.cxx:346-437);
      4. reaction potential at the charges from the panel charges
         si area / e1 (bio/main.cxx:445-501), energy per charge
         0.5 * 332.112 * q * potential (bio/main.cxx:502).
    Returns a dict with the total energy, per-charge energies, the solution
    si, the reaction field gradient at the charges, the panel geometry and the
    GMRES history."""
    from .formats import panels

    c, area, nrm = panels(vert, face)
    xq = np.ascontiguousarray(xq, np.float32)
    q = np.asarray(q, np.float32)
    f = FMM(p=p, precision=precision, convention="exact")
    f.set_points(c, xq, ni_normals=nrm)
    vi, _ = f.evaluate(q, neq=GNI)
    un = -4 * np.pi * vi / e1
    op = InducedCharge(c, nrm, area, e1, e2, p=p, precision=precision)
    si, hist, it = gmres(op.matvec, un, precond=op.precondition, rtol=rtol, restart=restart, maxiter=maxiter)
    op.close()
    f.set_points(xq, c)
    vr, gr = f.evaluate((si * area / e1).astype(np.float32), neq=G)
    f.close()
    per = 0.5 * KCAL * q.astype(np.float64) * vr
    return {"energy": float(per.sum()), "per_charge": per, "si": si, "grad": gr, "centroids": c, "area": area,
            "normals": nrm, "history": hist, "iterations": it}


def solvation_energy_6node(mesh: Mesh, xq, q, e1=4.0, e2=20.0, p=10, precision="mixed", rtol=1e-5, restart=30,
                           maxiter=1000):
    """The 6-node pipeline of biobem/main.cxx:96-200: right-hand side
    un = -4 pi Gni(charges -> nodes, normals -vna) / e1, GMRES on Collocation,
    reaction potential at the charges from the quadrature points with
    strengths 0.5 hsg (sum_k phi_k si_k) / e1, energy 0.5 * 332.112 * q * v."""
    xq = np.ascontiguousarray(xq, np.float32)
    q = np.asarray(q, np.float32)
    nodes = np.ascontiguousarray(mesh.nodes, np.float32)
    f = FMM(p=p, precision=precision, convention="exact")
    f.set_points(nodes, xq, ni_normals=np.ascontiguousarray(-mesh.normals, np.float32))
    vi, _ = f.evaluate(q, neq=GNI)
    un = -4 * np.pi * vi / e1
    op = Collocation(mesh, e1=e1, e2=e2, p=p, precision=precision)
    si, hist, it = gmres(op.matvec, un, precond=op.precondition, rtol=rtol, restart=restart, maxiter=maxiter)
    op.close()
    f.set_points(xq, np.ascontiguousarray(mesh.qx, np.float32))
    vr, _ = f.evaluate((mesh.qw * mesh.interpolate(si) / e1).astype(np.float32), neq=G, gradient=False)
    f.close()
    per = 0.5 * KCAL * q.astype(np.float64) * vr
    return {"energy": float(per.sum()), "per_charge": per, "si": si, "history": hist, "iterations": it}


# ---- batched solves: many independent surfaces in one FMM -------------------------
def ellipsoid_surface(npanel, axes, bumps=(), seed=0):
    """Triangulated ellipsoid with about npanel flat panels: the convex hull of
    a Fibonacci lattice on the unit sphere (2 n - 4 triangles for n points),
    mapped to semi-axes `axes` and scaled radially by 1 + sum of Gaussian
    bumps (centre on the unit sphere, amplitude, width).  Faces are listed
    counter-clockwise from outside, 1-based, as the reference's .face files.
    Returns (vertices float32 (n, 3), faces (m, 3))."""
    from scipy.spatial import ConvexHull

    n = max(8, npanel // 2 + 2)
    k = np.arange(n) + 0.5
    z = 1 - 2 * k / n
    r = np.sqrt(1 - z * z)
    ph = np.pi * (1 + 5 ** 0.5) * k
    u = np.stack([r * np.cos(ph), r * np.sin(ph), z], 1)
    tri = ConvexHull(u).simplices.copy()
    scale = np.ones(n)
    for c, amp, w in bumps:
        scale += amp * np.exp(-np.sum((u - np.asarray(c)) ** 2, 1) / (2 * w * w))
    v = u * np.asarray(axes, float) * scale[:, None]
    a, b, c = v[tri[:, 0]], v[tri[:, 1]], v[tri[:, 2]]
    out = np.einsum("ij,ij->i", np.cross(b - a, c - a), a + b + c) > 0
    tri[~out] = tri[~out][:, [0, 2, 1]]
    return v.astype(np.float32), tri + 1


class BatchedInducedCharge:
    """The operator of bio/main.cxx:27-53 for many independent surfaces at
    once: surface s is moved to cell s of the 2^G grid of the root cube and
    the FMM's group level G keeps the cells apart, so one evaluation (one set
    of launches over all panels) applies every surface's operator.  Vectors
    are concatenated in surface order; seg[s]:seg[s+1] is surface s."""

    def __init__(self, surfaces, e1=4.0, e2=20.0, p=8, precision="mixed", lmax=0, extra_levels=4, device=None,
                 tune=False, graph=True):
        import torch

        ns = len(surfaces)
        G = max(1, int(np.ceil(np.log2(max(2, ns)) / 3)))
        cells = 1 << G
        self.dev = torch.device(device or "cuda")  # the panel vectors live on the GPU
        # panel geometry of every surface at once on the device, the float32
        # arithmetic of formats.panels (bio/main.cxx:257-281)
        vs = [np.asarray(v, np.float32) for v, _ in surfaces]
        voff = np.cumsum([0] + [len(v) for v in vs[:-1]])
        V = torch.from_numpy(np.concatenate(vs)).to(self.dev)
        Fi = torch.from_numpy(np.concatenate([np.asarray(f, np.int64) - 1 + o for (_, f), o in zip(surfaces, voff)]))
        v = V[Fi.to(self.dev)]
        c = v[:, 0] / 3.0 + v[:, 1] / 3.0 + v[:, 2] / 3.0
        nv = torch.cross(v[:, 2] - v[:, 0], v[:, 1] - v[:, 0], dim=1)
        area = (0.5 * torch.sqrt((nv.double() ** 2).sum(1))).float()
        nrm = nv / area[:, None] / 2.0
        seg = np.cumsum([0] + [len(f) for _, f in surfaces])
        ext = max(float(np.abs(x - x.mean(0)).max()) for x in vs)
        h = 2.2 * ext  # cell size: every surface strictly inside its cell
        self.G, self.seg = G, seg
        self.e1, self.e2 = float(e1), float(e2)
        self.n = int(seg[-1])
        seg_id = torch.from_numpy(np.repeat(np.arange(ns), np.diff(seg))).to(self.dev)
        # surface s moves to the centre of cell s: c - mean(c_s) + centre_s
        cnt = torch.from_numpy(np.diff(seg).astype(np.float64)).to(self.dev)
        mean = torch.zeros((ns, 3), dtype=torch.float64, device=self.dev).index_add_(0, seg_id, c.double())
        mean = (mean / cnt[:, None]).float()
        s_ = np.arange(ns)
        centre = torch.from_numpy(((np.stack([s_ % cells, (s_ // cells) % cells, s_ // (cells * cells)], 1) + 0.5)
                                   * h).astype(np.float64)).to(self.dev)
        self.area = area.double()
        self.seg_id = seg_id
        self.nseg = ns
        self.f = FMM(p=p, lmax=lmax or G + extra_levels, precision=precision, convention="exact",
                     box=(0.0, 0.0, 0.0, float(h * cells)))
        self.f.set_groups(G)
        self._c = ((c - mean[seg_id]).double() + centre[seg_id]).float().contiguous()
        self._n = nrm.contiguous()
        self.f.set_points(self._c, ni_normals=self._n)
        self.diag = (self.e1 + self.e2) / (2 * self.e1 * (self.e1 - self.e2))
        # the panels in each surface's own coordinates (device), for the
        # right-hand sides and energies
        self.pc, self.pa, self.pn = c.contiguous(), area, nrm.contiguous()
        # strengths buffer of the matvec; optionally the M2L launch shapes
        # tuned on this tree (pays off only over many matvecs: a tuning pass
        # costs about as much as 20 of them); one evaluation captured as a
        # HIP graph that every matvec replays
        self._q = (self.area / self.e1).float().contiguous()
        self.tuning = self.f.autotune(self._q, neq=GNI, reps=2, levels=[self.f.lmax], quick=True) if tune else None
        self._v = None
        if graph:
            self._v, _ = self.f.capture(self._q, neq=GNI, gradient=False)

    def matvec(self, x):
        if self._v is None:
            v, _ = self.f.evaluate((x * self.area / self.e1).float(), neq=GNI)
            return v + self.diag * x
        self._q.copy_(x * self.area / self.e1)
        self.f.replay()
        return self._v + self.diag * x

    def precondition(self, x):
        return x / (self.e1 + self.e2) * (2 * self.e1 * (self.e1 - self.e2))

    def close(self):
        self.f.close()


def gmres_batched(matvec, b, seg_id, nseg, precond=None, rtol=1e-6, restart=30, maxiter=300, chunk=1024):
    """Independent right-preconditioned restarted GMRES for every segment of
    a block-diagonal system, in lockstep: one matvec per iteration serves all
    segments; inner products, Givens rotations and updates are per segment
    (FP64).  b and the iterates are torch tensors (device-resident for a GPU
    matvec); segment s is the contiguous run of seg_id == s.

    Layout: the Krylov vectors are kept with every segment padded to whole
    chunks of `chunk` values (zeros in the padding), as a (chunks, restart+1,
    chunk) array, so that the orthogonalisation is block classical
    Gram-Schmidt with one reorthogonalisation pass: per pass, one batched
    product gives the inner products of the new vector with every basis
    vector chunk by chunk, a sum over each segment's chunks turns them into
    per-segment coefficients, and one batched product subtracts the
    projections.  Segment sums run over chunks (about n / chunk values), not
    over single entries.
    Host synchronisation: the host waits for iteration k-1's convergence flag
    only after queueing iteration k, so the GPU never idles on the host; the
    restart's true residual is the one blocking read per cycle.
    Returns (x, per-segment relative residuals, iterations)."""
    import torch

    dev, dt = b.device, torch.float64
    gpu = dev.type == "cuda"
    M = precond or (lambda v: v)
    n, C = b.shape[0], int(chunk)
    if n > 1 and not bool((seg_id[1:] >= seg_id[:-1]).all()):
        raise ValueError("gmres_batched needs every segment contiguous and in order (seg_id non-decreasing)")
    counts = torch.bincount(seg_id, minlength=nseg).cpu().numpy().astype(np.int64)
    nch = (counts + C - 1) // C
    nc = int(nch.sum())
    cstart = np.concatenate([[0], np.cumsum(nch)])[:-1]
    start = np.concatenate([[0], np.cumsum(counts)])[:-1]
    chunk_seg = torch.from_numpy(np.repeat(np.arange(nseg), nch)).to(dev)
    # padded position of every entry (segments are contiguous and in order)
    sid = seg_id.long()
    pos = (torch.from_numpy(cstart * C - start).to(dev)[sid] + torch.arange(n, device=dev)).contiguous()

    def pack(v):
        out = torch.zeros(nc * C, dtype=dt, device=dev)
        out[pos] = v
        return out.view(nc, C)

    def unpack(v):
        return v.reshape(-1)[pos]

    def segsum(cv):  # per-chunk values (nc, ...) -> per-segment sums (nseg, ...)
        return torch.zeros((nseg,) + tuple(cv.shape[1:]), dtype=dt, device=dev).index_add_(0, chunk_seg, cv)

    def snorm(v):  # per-segment 2-norms of a packed vector
        return torch.sqrt(segsum((v * v).sum(1)))

    def safe(a):
        return torch.where(a > 0, a, torch.ones_like(a))

    bp = pack(b.to(dt))
    nb = safe(snorm(bp))
    x = torch.zeros(n, dtype=dt, device=dev)
    it = 0
    r = pack(b - matvec(x))
    beta = snorm(r)
    res = beta / nb
    V = None
    while bool((res > rtol).any()) and it < maxiter:
        m = min(restart, maxiter - it)
        if V is None or V.shape[1] < m + 1:
            V = torch.empty(nc, m + 1, C, dtype=dt, device=dev)
        V[:, 0, :] = r / safe(beta)[chunk_seg][:, None]
        H = torch.zeros(nseg, m + 1, m, dtype=dt, device=dev)
        g = torch.zeros(nseg, m + 1, dtype=dt, device=dev)
        g[:, 0] = beta
        cs = torch.zeros(nseg, m, dtype=dt, device=dev)
        sn = torch.zeros(nseg, m, dtype=dt, device=dev)
        kk = 0
        prev = None  # (event, pinned flag) of the previous iteration
        for k in range(m):
            w = pack(matvec(M(unpack(V[:, k, :]))))
            Vk = V[:, :k + 1, :]
            for _ in range(2):
                h = segsum(torch.bmm(Vk, w.unsqueeze(2))[:, :, 0])  # (nseg, k+1)
                H[:, :k + 1, k] += h
                w = w - torch.bmm(Vk.transpose(1, 2), h[chunk_seg].unsqueeze(2))[:, :, 0]
            hn = snorm(w)
            H[:, k + 1, k] = hn
            V[:, k + 1, :] = w / safe(hn)[chunk_seg][:, None]
            for i in range(k):
                t = cs[:, i] * H[:, i, k] + sn[:, i] * H[:, i + 1, k]
                H[:, i + 1, k] = -sn[:, i] * H[:, i, k] + cs[:, i] * H[:, i + 1, k]
                H[:, i, k] = t
            d = torch.sqrt(H[:, k, k] ** 2 + H[:, k + 1, k] ** 2)
            d1 = safe(d)
            cs[:, k] = torch.where(d > 0, H[:, k, k] / d1, torch.ones_like(d))
            sn[:, k] = torch.where(d > 0, H[:, k + 1, k] / d1, torch.zeros_like(d))
            H[:, k, k] = d
            H[:, k + 1, k] = 0
            g[:, k + 1] = -sn[:, k] * g[:, k]
            g[:, k] = cs[:, k] * g[:, k]
            it += 1
            kk = k + 1
            done = (g[:, k + 1].abs() / nb <= rtol).all()
            if not gpu:
                if bool(done):
                    break
                continue
            flag = torch.empty((), dtype=torch.bool, pin_memory=True)
            flag.copy_(done, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record()
            # iteration k is queued: now wait for iteration k-1's flag (the
            # GPU meanwhile runs iteration k)
            if prev is not None:
                prev[0].synchronize()
                if bool(prev[1]):
                    break
            prev = (ev, flag)
        R = H[:, :kk, :kk]
        diag = torch.diagonal(R, dim1=1, dim2=2)
        R = R + torch.diag_embed(torch.where(diag == 0, torch.ones_like(diag), torch.zeros_like(diag)))
        y = torch.linalg.solve_triangular(R, g[:, :kk, None], upper=True)[:, :, 0]
        upd = torch.bmm(V[:, :kk, :].transpose(1, 2), y[chunk_seg].unsqueeze(2))[:, :, 0]
        x = x + M(unpack(upd))
        r = pack(b - matvec(x))
        beta = snorm(r)
        res = beta / nb
    return x, res, it


def batched_solvation(surfaces, charges, e1=4.0, e2=20.0, p=8, precision="mixed", rtol=1e-6, restart=30,
                      maxiter=300, extra_levels=4, normalize="size", tune=False, graph=True):
    """Solve the induced-charge equation of bio/main.cxx for every surface of
    a batch at once (each with its own point charges inside) and return the
    per-surface solvation energies (kcal/mol), the solution and the GMRES
    history.  charges[s] = (positions (m, 3), charges (m,)) in the
    surface's own coordinates.

    normalize: scale every surface (and its charges) by a_s.  The operator
    (area times the normal-derivative kernel) is scale-free, the right-hand
    side scales as 1/a^2 and the reaction potential as 1/a, so the energy of
    the scaled problem times a_s is the surface's energy; the returned si is
    the scaled problem's.  "size" (or True): every surface fills its cell of
    the shared tree.  "density": a surface of n_s panels is scaled to
    sqrt(n_s / n_max) of the cell size, so every surface has about the
    same number of panels per leaf box (a small surface spans fewer boxes,
    a shallower effective tree).  False: as given.
    Returns also times_s: operator set-up, right-hand sides, GMRES and
    energies (seconds, each closed by a device synchronisation)."""
    import torch

    scale = np.ones(len(surfaces))
    if normalize:
        ext = [float(np.abs(np.asarray(v) - np.asarray(v).mean(0)).max()) for v, _ in surfaces]
        scale = np.max(ext) / np.array(ext)
        if normalize == "density":
            npan = np.array([len(f) for _, f in surfaces], float)
            scale = scale * np.sqrt(npan / npan.max())
        surfaces = [((np.asarray(v, np.float64) * a).astype(np.float32), f) for (v, f), a in zip(surfaces, scale)]
        charges = [((np.asarray(x, np.float64) * a).astype(np.float32), q) for (x, q), a in zip(charges, scale)]
    import time

    def now():
        torch.cuda.synchronize()
        return time.perf_counter()

    t0 = now()
    op = BatchedInducedCharge(surfaces, e1, e2, p=p, precision=precision, extra_levels=extra_levels, tune=tune,
                              graph=graph)
    t1 = now()
    # right-hand side: -4 pi Gni(charges -> centroids) / e1 per surface, on
    # the oracle-free GPU direct sum of the engine
    f = FMM(p=p, precision="fp64", convention="exact")
    b = torch.zeros(op.n, dtype=torch.float64, device=op.dev)
    energies = []
    dq = [(torch.from_numpy(np.asarray(xq, np.float32)).to(op.dev), torch.from_numpy(np.asarray(q, np.float32)).to(op.dev))
          for xq, q in charges]
    for s, (xq, q) in enumerate(dq):
        sl = slice(int(op.seg[s]), int(op.seg[s + 1]))
        v, _ = f.direct(op.pc[sl], xq, q, neq=GNI, ni_normals=op.pn[sl])
        b[sl] = -4 * np.pi * v / e1
    t2 = now()
    x, res, it = gmres_batched(op.matvec, b, op.seg_id, op.nseg, precond=op.precondition, rtol=rtol,
                               restart=restart, maxiter=maxiter)
    t3 = now()
    si = x.cpu().numpy()
    for s, (xq, q) in enumerate(dq):
        sl = slice(int(op.seg[s]), int(op.seg[s + 1]))
        vr, _ = f.direct(xq, op.pc[sl], (x[sl] * op.pa[sl] / e1).float(), neq=G)
        energies.append(float(0.5 * KCAL * (q.double() * vr).sum()) * scale[s])
    t4 = now()
    f.close()
    tuning = op.tuning
    op.close()
    return {"energies": np.array(energies), "si": si, "residuals": res.cpu().numpy(), "iterations": it,
            "seg": op.seg, "scale": scale, "tuning": tuning,
            "times_s": {"operator_setup": t1 - t0, "rhs": t2 - t1, "gmres": t3 - t2, "energies": t4 - t3}}
