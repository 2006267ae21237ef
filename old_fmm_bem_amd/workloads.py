"""Synthetic inputs of the matvec's configurations, shared by bench.py and the
full-size GPU tests so that both measure the same problems:

  cube:        N points uniform in [0,1)^3, charges uniform in [-1,1]
               (the headline workload, and the reference's test/test.cxx
               distribution on the unit cube);
  sphere:      N points uniform on the unit sphere with complex charges
               (the oscillatory Helmholtz workload);
  plummer:     a Plummer sphere (scale 1, truncated at r <= rmax) with
               charges uniform in [0,1] (the adjoint workload);
  bem_sphere:  the 3 Gauss points of each triangle of an icosphere with
               unit normals and a smooth random spherical-harmonic density
               times the quadrature weights (the BEM matvec workload);
  ellipsoids:  the batched-solve workload: ellipsoids with Gaussian bumps,
               log-uniform panel counts, three charges inside each.
"""
from __future__ import annotations

import os

import numpy as np


def cube(n, seed=0):
    rng = np.random.default_rng(seed)
    x = rng.random((n, 3), dtype=np.float32)
    q = (2 * rng.random(n, dtype=np.float32) - 1).astype(np.float32)
    return x, q


def sphere(n, seed=0):
    """Points uniform on the unit sphere, complex charges with parts in [-1,1]."""
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, 3))
    x = (x / np.linalg.norm(x, axis=1)[:, None]).astype(np.float32)
    q = ((2 * rng.random(n) - 1) + 1j * (2 * rng.random(n) - 1)).astype(np.complex64)
    return x, q


def plummer(n, seed=0, rmax=10.0):
    """Plummer sphere of scale 1 truncated at r <= rmax (radius from the
    inverse cumulative mass, isotropic direction), charges uniform in [0,1)."""
    rng = np.random.default_rng(seed)
    out = np.zeros((0, 3))
    while len(out) < n:
        m = rng.random(2 * n)
        r = 1.0 / np.sqrt(np.maximum(m, 1e-12) ** (-2.0 / 3.0) - 1.0)
        v = rng.standard_normal((2 * n, 3))
        v /= np.linalg.norm(v, axis=1)[:, None]
        out = np.concatenate([out, (r[:, None] * v)[r <= rmax]])
    return out[:n].astype(np.float32), rng.random(n).astype(np.float32)


def bem_sphere(subdiv=6, seed=0, degree=8):
    """Gauss points (3 per triangle) of the icosphere with 20*4^subdiv
    triangles, unit normals, and strengths = quadrature weight times a random
    real spherical-harmonic field of degree <= `degree` (coefficients N(0,1))."""
    from scipy.special import sph_harm_y

    from . import bem

    rng = np.random.default_rng(seed)
    m = bem.icosphere(subdiv)
    x = m.qx.astype(np.float32)
    nrm = (m.qx / np.linalg.norm(m.qx, axis=1)[:, None]).astype(np.float32)
    th = np.arccos(np.clip(m.qx[:, 2] / np.linalg.norm(m.qx, axis=1), -1, 1))
    ph = np.arctan2(m.qx[:, 1], m.qx[:, 0])
    sig = np.zeros(len(x))
    for deg in range(degree + 1):
        for o in range(deg + 1):
            y = sph_harm_y(deg, o, th, ph)
            sig += rng.standard_normal() * y.real + (rng.standard_normal() * y.imag if o else 0)
    return x, nrm, (m.qw * sig).astype(np.float32)


def _surface(args):
    from . import bem

    n, axes, bumps, seed = args
    return bem.ellipsoid_surface(n, axes, bumps, seed=seed)


def ellipsoids(ns=256, seed=0, nmin=2000, nmax=200000, workers=None):
    """ns ellipsoids (semi-axes uniform in [0.5, 2], 20 Gaussian bumps of
    amplitude <= 0.1), panel counts log-uniform in [nmin, nmax], three unit
    charges inside each.  Returns (surfaces [(vertices, faces)], charges
    [(positions, values)]).  The convex hulls are built in worker processes
    (a fresh interpreter each) when workers > 1."""
    rng = np.random.default_rng(seed)
    specs, charges = [], []
    for s in range(ns):
        n = int(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))
        axes = rng.uniform(0.5, 2.0, 3)
        bumps = [(u / np.linalg.norm(u), rng.uniform(0, 0.1), rng.uniform(0.15, 0.4))
                 for u in rng.standard_normal((20, 3))]
        specs.append((n, axes, bumps, s))
        charges.append(((rng.uniform(-0.3, 0.3, (3, 3)) * axes).astype(np.float32), np.ones(3, np.float32)))
    if workers is None:
        workers = min(16, os.cpu_count() or 1)
    if workers > 1 and ns > 1:
        import multiprocessing as mp
        from concurrent.futures import ProcessPoolExecutor

        with ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn")) as ex:
            surfaces = list(ex.map(_surface, specs, chunksize=4))
    else:
        surfaces = [_surface(a) for a in specs]
    return surfaces, charges
