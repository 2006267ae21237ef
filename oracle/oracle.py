"""ctypes wrapper around the oracle (liborc.so): the CPU FP64 restatement of
the reference Laplace FMM.

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py may import this module; it is the checker,
never the thing measured or shipped.

Conventions (see fmm_oracle.h): ``REF`` reproduces the reference arithmetic
(eps softening, eps in radii, float tables/pi, float outputs); ``EXACT`` drops
those quirks (unsoftened kernel, self pairs excluded).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

REF, EXACT = 0, 1
NEQ_G, NEQ_GNI, NEQ_GNJ = 0, 1, 2
NEBM = 189

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class _Opts(C.Structure):
    _fields_ = [("mp", C.c_int), ("lmax", C.c_int), ("neq", C.c_int), ("conv", C.c_int),
                ("nthreads", C.c_int), ("xmin", C.c_float), ("ymin", C.c_float),
                ("zmin", C.c_float), ("rd", C.c_float), ("fp64_out", C.c_int)]


def build() -> str:
    """Compile liborc.so in place (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE, "liborc.so"], check=True)
    return os.path.join(_HERE, "liborc.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liborc.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        fp, dp, ip = C.POINTER(C.c_float), C.POINTER(C.c_double), C.POINTER(C.c_int)
        L.orc_new.restype = C.c_void_p
        L.orc_new.argtypes = [C.POINTER(_Opts), C.c_int, fp, fp, fp, fp, fp, fp, C.c_int, fp, fp, fp, fp, fp, fp]
        L.orc_free.argtypes = [C.c_void_p]
        L.orc_run.argtypes = [C.c_void_p, fp, dp, dp, dp, dp]
        L.orc_times.argtypes = [C.c_void_p, dp]
        L.orc_info.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orc_perm.argtypes = [C.c_void_p, ip, ip]
        L.orc_boxes_i.argtypes = [C.c_void_p, C.c_int, ip, ip, ip]
        L.orc_boxes_j.argtypes = [C.c_void_p, C.c_int, ip, ip, ip]
        L.orc_list.argtypes = [C.c_void_p, C.c_int, C.c_int, ip, ip]
        L.orc_bx.argtypes = [C.c_void_p, dp]
        L.orc_ax.argtypes = [C.c_void_p, dp]
        L.orc_ax_level.argtypes = [C.c_void_p, C.c_int, dp]
        L.orc_precalc.argtypes = [C.c_int, C.c_int, dp, dp, dp, dp, dp]
        L.orc_direct.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, fp, fp, fp, fp, fp, fp,
                                 C.c_int, fp, fp, fp, fp, fp, fp, fp, dp, dp, dp, dp]
        L.orc_direct_yukawa.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, fp, fp, fp, C.c_int, fp, fp, fp, fp,
                                        dp, dp, dp, dp]
        L.orc_direct_helmholtz.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, fp, fp, fp, C.c_int, fp, fp, fp,
                                           fp, fp, dp, dp]
        L.orc_p2m_box.argtypes = [C.c_int, C.c_int, C.c_int, fp, fp, fp, fp, C.c_double, C.c_double,
                                  C.c_double, dp]
        L.orc_l2p_box.argtypes = [C.c_int, C.c_int, C.c_int, fp, fp, fp, C.c_double, C.c_double,
                                  C.c_double, dp, dp, dp, dp, dp]
        L.orc_m2l_pair.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, dp, dp]
        _LIB = L
    return _LIB


def _f(a):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a


def _pf(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


def _pd(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def _pi(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_int))


class OracleFMM:
    """Tree + lists built once (as the reference's setup chain), then ``run``."""

    def __init__(self, xi, xj, *, mp=10, lmax=3, neq=NEQ_G, conv=REF, box=None, ni_normals=None,
                 nj_normals=None, nthreads=0, fp64_out=False):
        xi = np.ascontiguousarray(xi, dtype=np.float32)
        xj = np.ascontiguousarray(xj, dtype=np.float32)
        if box is None:
            box = bounding_cube(np.concatenate([xi, xj]))
        self.box = tuple(np.float32(b) for b in box)
        o = _Opts(mp, lmax, neq, conv, nthreads, *self.box, int(fp64_out))
        self.mp, self.lmax, self.neq, self.conv = mp, lmax, neq, conv
        self.ni, self.nj = len(xi), len(xj)
        self._keep = [_f(xi[:, k]) for k in range(3)] + [_f(xj[:, k]) for k in range(3)]
        ni3 = [None] * 3 if ni_normals is None else [_f(ni_normals[:, k]) for k in range(3)]
        nj3 = [None] * 3 if nj_normals is None else [_f(nj_normals[:, k]) for k in range(3)]
        self._keep += ni3 + nj3
        k = self._keep
        self.h = lib().orc_new(C.byref(o), self.ni, _pf(k[0]), _pf(k[1]), _pf(k[2]), *map(_pf, ni3),
                               self.nj, _pf(k[3]), _pf(k[4]), _pf(k[5]), *map(_pf, nj3))
        if not self.h:
            raise ValueError("orc_new rejected the options")

    def close(self):
        if getattr(self, "h", None):
            lib().orc_free(self.h)
            self.h = None

    __del__ = close

    def run(self, q):
        q = _f(q)
        v = np.zeros(self.ni)
        g = np.zeros((3, self.ni))
        lib().orc_run(self.h, _pf(q), _pd(v), _pd(g[0]), _pd(g[1]), _pd(g[2]))
        return v, g.T.copy()

    def times(self):
        t = np.zeros(8)
        lib().orc_times(self.h, _pd(t))
        return dict(zip(["p2p", "p2m", "m2m", "m2l", "l2l", "l2p", "precalc", "tree"], t))

    def info(self, what, lev=0):
        return lib().orc_info(self.h, what, lev)

    def perm(self):
        a = np.zeros(self.ni, np.int32)
        b = np.zeros(self.nj, np.int32)
        lib().orc_perm(self.h, _pi(a), _pi(b))
        return a, b

    def boxes(self, side, lev):
        n = self.info(3 if side == "i" else 4, lev)
        k = np.zeros(n, np.int32)
        f = np.zeros(n, np.int32)
        l = np.zeros(n, np.int32)
        fn = lib().orc_boxes_i if side == "i" else lib().orc_boxes_j
        fn(self.h, lev, _pi(k), _pi(f), _pi(l))
        return k, f, l

    def interaction_list(self, lev, kind):
        n = self.info(3, lev)
        nij = np.zeros(n, np.int32)
        neij = np.zeros(n * NEBM, np.int32)
        lib().orc_list(self.h, lev, kind, _pi(nij), _pi(neij))
        neij = neij.reshape(n, NEBM)
        return nij, np.concatenate([neij[i, :nij[i]] for i in range(n)]) if n else np.zeros(0, np.int32)

    def bx(self):
        nmp = self.info(2)
        a = np.zeros(self.info(1) * nmp * 2)
        lib().orc_bx(self.h, _pd(a))
        return a.view(np.complex128).reshape(-1, nmp)

    def ax(self):
        nmp = self.info(2)
        a = np.zeros(self.info(3, self.lmax) * nmp * 2)
        lib().orc_ax(self.h, _pd(a))
        return a.view(np.complex128).reshape(-1, nmp)

    def ax_level(self, lev):
        nmp = self.info(2)
        a = np.zeros(self.info(3, lev) * nmp * 2)
        lib().orc_ax_level(self.h, lev, _pd(a))
        return a.view(np.complex128).reshape(-1, nmp)


def bounding_cube(x):
    """Root cube as the BEM caller builds it (bio/main.cxx:355-368): float
    min corner; per axis the float extent xmax-xmin times the double literal
    1+1e-5, the double product rounded once into the float rdx; rd the
    largest of rdx, rdy, rdz."""
    x = np.asarray(x, dtype=np.float32)
    lo = x.min(axis=0).astype(np.float32)
    hi = x.max(axis=0).astype(np.float32)
    ext = (hi - lo).astype(np.float32)  # float - float
    rdk = (ext.astype(np.float64) * (1 + 1e-5)).astype(np.float32)  # double product -> float
    rd = np.float32(max(rdk[0], rdk[1], rdk[2]))
    return float(lo[0]), float(lo[1]), float(lo[2]), float(rd)


def fmm(xi, xj, q, **kw):
    f = OracleFMM(xi, xj, **kw)
    try:
        return f.run(q)
    finally:
        f.close()


def direct(xi, xj, q, *, neq=NEQ_G, conv=REF, ni_normals=None, nj_normals=None, nthreads=0):
    xi = np.ascontiguousarray(xi, dtype=np.float32)
    xj = np.ascontiguousarray(xj, dtype=np.float32)
    q = _f(q)
    ni, nj = len(xi), len(xj)
    a = [_f(xi[:, k]) for k in range(3)]
    b = [_f(xj[:, k]) for k in range(3)]
    n1 = [None] * 3 if ni_normals is None else [_f(ni_normals[:, k]) for k in range(3)]
    n2 = [None] * 3 if nj_normals is None else [_f(nj_normals[:, k]) for k in range(3)]
    v = np.zeros(ni)
    g = np.zeros((3, ni))
    lib().orc_direct(neq, conv, nthreads, ni, *map(_pf, a), *map(_pf, n1), nj, *map(_pf, b), *map(_pf, n2),
                     _pf(q), _pd(v), _pd(g[0]), _pd(g[1]), _pd(g[2]))
    return v, g.T.copy()


def direct_yukawa(xi, xj, q, lam, *, conv=REF, nthreads=0):
    """kernel/Helmholtz_G_dir.cxx: potential (and, EXACT only, gradient)."""
    xi = np.ascontiguousarray(xi, dtype=np.float32)
    xj = np.ascontiguousarray(xj, dtype=np.float32)
    q = _f(q)
    ni, nj = len(xi), len(xj)
    a = [_f(xi[:, k]) for k in range(3)]
    b = [_f(xj[:, k]) for k in range(3)]
    v = np.zeros(ni)
    g = np.zeros((3, ni))
    lib().orc_direct_yukawa(conv, nthreads, float(lam), ni, *map(_pf, a), nj, *map(_pf, b), _pf(q), _pd(v),
                            _pd(g[0]), _pd(g[1]), _pd(g[2]))
    return v, g.T.copy()


def direct_helmholtz(xi, xj, q, k, *, conv=EXACT, nthreads=0):
    """sum_j q_j exp(i k r) / (4 pi r), complex q: FP64 from float inputs."""
    xi = np.ascontiguousarray(xi, dtype=np.float32)
    xj = np.ascontiguousarray(xj, dtype=np.float32)
    q = np.asarray(q)
    qr, qi = _f(q.real), _f(q.imag)
    ni, nj = len(xi), len(xj)
    a = [_f(xi[:, c]) for c in range(3)]
    b = [_f(xj[:, c]) for c in range(3)]
    vr, vi = np.zeros(ni), np.zeros(ni)
    lib().orc_direct_helmholtz(conv, nthreads, float(k), ni, *map(_pf, a), nj, *map(_pf, b), _pf(qr), _pf(qi),
                               _pd(vr), _pd(vi))
    return vr + 1j * vi


def precalc(mp, conv=REF):
    mp2 = mp * mp
    fac = np.zeros(4 * mp2)
    anm = np.zeros(4 * mp2)
    sr = np.zeros(mp2 * mp2)
    ynm = np.zeros(8 * mp2)
    dnm = np.zeros(2 * 1024 * mp * mp2)
    lib().orc_precalc(mp, conv, _pd(fac), _pd(anm), _pd(sr), _pd(ynm), _pd(dnm))
    return dict(fac=fac, anm=anm, sr=sr, ynm=ynm.view(np.complex128),
                dnm=dnm.view(np.complex128).reshape(1024, mp, mp2))


def p2m_box(mp, conv, x, q, c):
    x = np.ascontiguousarray(x, dtype=np.float32)
    nmp = mp * (mp + 1) // 2
    m = np.zeros(2 * nmp)
    lib().orc_p2m_box(mp, conv, len(x), *[_pf(_f(x[:, k])) for k in range(3)], _pf(_f(q)),
                      float(c[0]), float(c[1]), float(c[2]), _pd(m))
    return m.view(np.complex128)


def l2p_box(mp, conv, x, c, L):
    x = np.ascontiguousarray(x, dtype=np.float32)
    n = len(x)
    L = np.ascontiguousarray(L, dtype=np.complex128)
    v = np.zeros(n)
    g = np.zeros((3, n))
    lib().orc_l2p_box(mp, conv, n, *[_pf(_f(x[:, k])) for k in range(3)], float(c[0]), float(c[1]),
                      float(c[2]), _pd(L.view(np.float64)), _pd(v), _pd(g[0]), _pd(g[1]), _pd(g[2]))
    return v, g.T.copy()


def m2l_pair(mp, conv, d, rb, M):
    M = np.ascontiguousarray(M, dtype=np.complex128)
    L = np.zeros_like(M)
    lib().orc_m2l_pair(mp, conv, int(d[0]), int(d[1]), int(d[2]), float(rb), _pd(M.view(np.float64)),
                       _pd(L.view(np.float64)))
    return L
