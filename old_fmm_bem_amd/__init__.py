"""MI355X FMM engine: Python host layer over the C ABI of libfmmx.so.

Mirrors the reference's FMM matvec interface (fmm/fmm.cxx:29,
``fmm(mi, mj, neq)`` with neq 0 = G, 1 = Gni, 2 = Gnj) as an explicit object:

    f = FMM(p=10, lmax=4, precision="mixed", convention="exact")
    f.set_points(xi, xj)               # tree + lists (sorti/sortj/boxallocate/...)
    v, g = f.evaluate(q, neq=0)        # one matvec

Arrays may be numpy (host) or torch CUDA tensors (device pointers; then the
outputs are torch tensors on the same device).  The compute always runs in
libfmmx.so's HIP kernels; there is no CPU fallback -- a missing or unloadable
library raises.
"""
from __future__ import annotations

import ctypes as C
import os
import time

import numpy as np

__all__ = ["FMM", "FMMError", "lib", "operator", "auto_lmax", "G", "GNI", "GNJ", "YUKAWA", "HELMHOLTZ", "MIXED", "FP64", "REF",
           "EXACT"]

G, GNI, GNJ, YUKAWA, HELMHOLTZ = 0, 1, 2, 10, 20
MIXED, FP64 = 0, 1
REF, EXACT = 0, 1
NEBM = 189

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_PREC = {"mixed": MIXED, "fp32": MIXED, "fp64": FP64, MIXED: MIXED, FP64: FP64}
_CONV = {"ref": REF, "exact": EXACT, REF: REF, EXACT: EXACT}


class FMMError(RuntimeError):
    pass


class _Config(C.Structure):
    _fields_ = [("p", C.c_int), ("lmax", C.c_int), ("precision", C.c_int), ("convention", C.c_int),
                ("xmin", C.c_float), ("ymin", C.c_float), ("zmin", C.c_float), ("rd", C.c_float)]


class Stats(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("ms_tree", "ms_lists", "ms_p2p", "ms_p2m", "ms_m2m", "ms_m2l",
                                          "ms_l2l", "ms_l2p", "ms_total")] + \
               [(n, C.c_int) for n in ("lmax", "nbi_leaf", "nbj_leaf", "nbi_total", "nbj_total")] + \
               [(n, C.c_int64) for n in ("p2p_pairs", "m2l_pairs", "m2l_tiles", "m2l_pairs_leaf")] + \
               [("ms_m2l_leaf", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_SYMBOLS = {
    "fmmx_create": (C.c_int, [C.POINTER(_Config), C.c_int, C.c_void_p, C.POINTER(C.c_void_p)]),
    "fmmx_destroy": (C.c_int, [C.c_void_p]),
    "fmmx_set_points": (C.c_int, [C.c_void_p, C.c_int] + [C.c_void_p] * 6 + [C.c_int] + [C.c_void_p] * 6 + [C.c_int]),
    "fmmx_evaluate": (C.c_int, [C.c_void_p, C.c_int] + [C.c_void_p] * 5 + [C.c_int]),
    "fmmx_direct": (C.c_int, [C.c_void_p, C.c_int, C.c_int] + [C.c_void_p] * 6 + [C.c_int] + [C.c_void_p] * 11
                    + [C.c_int]),
    "fmmx_set_helmholtz": (C.c_int, [C.c_void_p, C.c_double]),
    "fmmx_set_groups": (C.c_int, [C.c_void_p, C.c_int]),
    "fmmx_graph_capture": (C.c_int, [C.c_void_p, C.c_int] + [C.c_void_p] * 5),
    "fmmx_graph_launch": (C.c_int, [C.c_void_p]),
    "fmmx_set_m2l_variant": (C.c_int, [C.c_void_p] + [C.c_int] * 5),
    "fmmx_get_m2l_variant": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "fmmx_set_yukawa": (C.c_int, [C.c_void_p, C.c_double]),
    "fmmx_set_domain": (C.c_int, [C.c_void_p, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float]),
    "fmmx_sort": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                            C.c_int]),
    "fmmx_sort_noalloc": (C.c_int, [C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                    C.c_float, C.c_float, C.c_float, C.c_float, C.c_void_p, C.c_void_p, C.c_void_p]),
    "fmmx_synchronize": (C.c_int, [C.c_void_p]),
    "fmmx_get_stats": (C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    "fmmx_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "fmmx_get_count": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "fmmx_get_domain": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "fmmx_auto_lmax": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p]),
    "fmmx_get_occupancy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "fmmx_evaluate_phase": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int] + [C.c_void_p] * 5 + [C.c_int]),
    "fmmx_get_multipoles_level": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "fmmx_evaluate_transpose": (C.c_int, [C.c_void_p, C.c_int] + [C.c_void_p] * 5 + [C.c_int]),
    "fmmx_gradients": (C.c_int, [C.c_void_p, C.c_int] + [C.c_void_p] * 5 + [C.c_int]),
    "fmmx_energy": (C.c_int, [C.c_void_p, C.c_int] + [C.c_void_p] * 4 + [C.c_int]),
    "fmmx_get_perm": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "fmmx_get_boxes": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "fmmx_get_list": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]),
    "fmmx_get_multipoles": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fmmx_get_locals": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "fmmx_operator": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_void_p]),
    "fmmx_last_error": (C.c_char_p, []),
    "fmmx_abi_version": (C.c_int, []),
}


def library_path() -> str:
    return os.path.join(_HERE, "libfmmx.so")


def auto_lmax(boxes, sumsq, p, precision="mixed"):
    """The engine's depth rule for lmax 0 in the EXACT convention: the leaf
    level 2..8 of least modelled cost, from the box statistics of a point set
    at levels 0..8 of its root cube (boxes[l] holding a point, sumsq[l] the
    sum of squared points per box)."""
    b = np.ascontiguousarray(boxes, dtype=np.int64)
    q = np.ascontiguousarray(sumsq, dtype=np.float64)
    if b.shape != (9,) or q.shape != (9,):
        raise ValueError("boxes and sumsq need 9 entries (levels 0..8)")
    lm = C.c_int(0)
    _check(lib().fmmx_auto_lmax(b.ctypes.data, q.ctypes.data, int(p), _PREC[precision], C.byref(lm)))
    return lm.value


def lib():
    """Load libfmmx.so (built by old_fmm_bem_amd/build.py); raises if absent."""
    global _LIB
    if _LIB is None:
        path = library_path()
        if not os.path.exists(path):
            raise FMMError(f"{path} is missing: build it with `python -m old_fmm_bem_amd.build` "
                           "(there is no CPU fallback)")
        L = C.CDLL(path)
        for name, (res, args) in _SYMBOLS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


_TORCH_READY = False


def _torch_first():
    """PyTorch's ROCm wheel carries its own HIP runtime.  When this library's
    runtime (the system one) opens the GPU first, torch's then reports no
    device; the other order works for both.  So torch, when installed, gets
    the device before the first handle is created."""
    global _TORCH_READY
    if not _TORCH_READY:
        _TORCH_READY = True
        try:
            import torch
        except ImportError:
            return
        torch.cuda.is_available()


def _check(rc):
    if rc != 0:
        raise FMMError(f"libfmmx error {rc}: {lib().fmmx_last_error().decode()}")


def _is_torch(a):
    return type(a).__module__.startswith("torch")


class _Arr:
    """Normalises an input column: numpy -> contiguous float32 host array,
    torch CUDA tensor -> contiguous float32 device tensor."""

    def __init__(self, a, device):
        self.dev = device
        if a is None:
            self.a, self.ptr = None, None
        elif device:
            import torch
            self.a = a.detach().to(torch.float32).contiguous()
            self.ptr = self.a.data_ptr()
        else:
            self.a = np.ascontiguousarray(a, dtype=np.float32)
            self.ptr = self.a.ctypes.data


def operator(p, convention, kind, a0, a1=0, a2=0, rb=1.0):
    """Dense real operator the engine applies (host only, no GPU): kind 0 M2M
    (octant a0), 1 L2L (octant a0), 2 M2L (offset a0,a1,a2 at box size rb);
    3-5 the same for YUKAWA (rb = lambda * box size), 6-8 for HELMHOLTZ
    (rb = k * box size, (2 p^2) x (2 p^2) on split expansions)."""
    d = 2 * p * p if kind in (6, 7, 8) else p * p
    A = np.zeros((d, d))
    _check(lib().fmmx_operator(p, _CONV[convention], kind, a0, a1, a2, float(rb), A.ctypes.data))
    return A


class FMM:
    def __init__(self, p=10, lmax=0, precision="mixed", convention="exact", box=None, device=0, stream=None,
                 lam=None, k=None):
        cfg = _Config(p, lmax, _PREC[precision], _CONV[convention], 0.0, 0.0, 0.0, 0.0)
        if box is not None:
            cfg.xmin, cfg.ymin, cfg.zmin, cfg.rd = [float(np.float32(b)) for b in box]
        self.p, self.lmax_cfg = p, lmax
        self.precision, self.convention = _PREC[precision], _CONV[convention]
        h = C.c_void_p()
        _torch_first()
        _check(lib().fmmx_create(C.byref(cfg), device, stream, C.byref(h)))
        self.h = h
        self.ni = self.nj = 0
        self._device = False
        if lam is not None:
            self.set_yukawa(lam)
        if k is not None:
            self.set_helmholtz(k)

    def capture(self, q, neq=G, gradient=True):
        """Capture one evaluation on device tensors as a HIP graph; returns the
        output tensors (v, g) that every replay() rewrites from the current
        contents of q (a float32 CUDA tensor the caller updates in place)."""
        import torch

        if not self._device:
            raise FMMError("capture needs device tensors (set_points with CUDA tensors)")
        want_g = gradient and neq in (G, YUKAWA)
        self._gq = q
        self._gv = torch.empty(self.ni, dtype=torch.float64, device=q.device)
        self._gg = torch.empty((3, self.ni), dtype=torch.float64, device=q.device) if want_g else None
        gp = [self._gg[k].data_ptr() for k in range(3)] if want_g else [None] * 3
        _check(lib().fmmx_graph_capture(self.h, neq, q.data_ptr(), self._gv.data_ptr(), *gp))
        return self._gv, (self._gg.T if want_g else None)

    def m2l_variant(self, lev):
        v = np.zeros(4, np.int32)
        _check(lib().fmmx_get_m2l_variant(self.h, lev, v.ctypes.data))
        return tuple(int(a) for a in v)

    def set_m2l_variant(self, lev, nt=0, staging=-1, wpg=0, nsplit=0):
        _check(lib().fmmx_set_m2l_variant(self.h, lev, nt, staging, wpg, nsplit))

    def autotune(self, q, neq=G, reps=3, margin=0.01, levels=None, quick=False):
        """Pick the M2L launch variant of every level by timing whole
        evaluations of q on the current points (so that the choice accounts
        for the other phases running beside the M2L): levels from the leaf
        up, each trying the MFMA staging/workgroup shapes and split counts
        around the default, the others held.  Returns {level: (n-tiles,
        staging, waves per workgroup, splits)}.  A candidate replaces the
        best so far only when faster by more than the fraction margin.
        levels: the levels to tune (default every level); quick: only the
        default shape and the FP16-split shapes, two split counts each."""
        import torch

        def timed():
            self.evaluate(q, neq=neq)
            self.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                self.evaluate(q, neq=neq)
            self.synchronize()
            return time.perf_counter() - t0

        del torch
        lmax = self.lmax
        shapes = [(1, 0, 8)] if self.precision == FP64 else \
            [(1, 0, 8), (1, 0, 4), (1, 1, 8)] + ([(1, 2, 16), (1, 2, 8)] if self.p <= 11 else [])
        if self.precision != FP64 and neq in (G, GNI, GNJ):
            # the error-compensated FP16 kernel (falls back to the FP32 shape
            # where it does not apply: Yukawa, REF convention, p outside 6..12)
            shapes += [(2, 5, 8), (2, 5, 4), (4, 5, 4), (1, 4, 8), (1, 3, 8), (1, 3, 4), (2, 3, 4)]
        if quick:
            shapes = [shapes[0]] + ([(2, 5, 8), (1, 4, 8)] if self.precision != FP64 and neq in (G, GNI, GNJ) else [])
        for lev in range(lmax, 1, -1):
            if self.nboxes("i", lev) == 0 or (levels is not None and lev not in levels):
                continue
            # the engine's default shape first: it stays unless another is
            # clearly faster
            self.set_m2l_variant(lev, 0, -1, 0, 0)
            dflt = tuple(self.m2l_variant(lev)[:3])
            best, tbest, seen = None, float("inf"), set()
            for nt, st, wpg in [dflt] + [sh for sh in shapes if sh != dflt]:
                self.set_m2l_variant(lev, nt, st, wpg, 0)
                base = self.m2l_variant(lev)[3]
                cand = {base, min(16, base * 2)} if quick else \
                    {base, max(1, base // 2), min(16, base * 2), max(1, base // 4), min(16, base * 4)}
                for ns in sorted(cand):
                    self.set_m2l_variant(lev, nt, st, wpg, ns)
                    eff = tuple(self.m2l_variant(lev))  # the launch that runs (shapes can coincide)
                    if eff in seen:
                        continue
                    seen.add(eff)
                    t = timed()
                    # the first (default) shape stays unless another is clearly
                    # faster: differences within the timing noise do not switch
                    if t < tbest * (1.0 - margin):
                        best, tbest = eff, t
            self.set_m2l_variant(lev, *best)
        return {lev: self.m2l_variant(lev) for lev in range(2, lmax + 1)}

    def replay(self):
        """Replay the captured evaluation (asynchronous on the handle's stream)."""
        _check(lib().fmmx_graph_launch(self.h))

    def set_groups(self, level):
        """Points in different boxes of this level are independent problems
        (0: off).  Call set_points afterwards."""
        _check(lib().fmmx_set_groups(self.h, int(level)))
        return self

    def set_helmholtz(self, k):
        """Wavenumber of the HELMHOLTZ kernel exp(i k r) / (4 pi r)."""
        _check(lib().fmmx_set_helmholtz(self.h, float(k)))
        return self

    def set_yukawa(self, lam):
        """Screening constant of the YUKAWA kernel (pi/2) exp(-lam r)/(lam r)."""
        _check(lib().fmmx_set_yukawa(self.h, float(lam)))
        return self

    def close(self):
        if getattr(self, "h", None):
            lib().fmmx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- set-up ---------------------------------------------------------
    def set_domain(self, lmax=0, box=None):
        """Replace the leaf level and root cube (xmin, ymin, zmin, rd) fixed at
        construction; box None: from the points at each set_points."""
        b = [0.0, 0.0, 0.0, 0.0] if box is None else [float(np.float32(v)) for v in box]
        _check(lib().fmmx_set_domain(self.h, int(lmax), *b))
        self.lmax_cfg = lmax
        return self

    def sort(self, x):
        """Morton keys at the leaf level and their stable sort, as
        sort/sorti.cxx:12-26: returns (keys in sorted order, perm) with
        perm[k] = index of the k-th point in key order."""
        n = int(x.shape[0])
        cols = [_Arr(np.asarray(x)[:, k], False) for k in range(3)]
        keys = np.zeros(max(n, 1), np.int32)
        perm = np.zeros(max(n, 1), np.int32)
        _check(lib().fmmx_sort(self.h, n, *[c.ptr for c in cols], keys.ctypes.data, perm.ctypes.data, 0))
        return keys[:n], perm[:n]

    def set_points(self, xi, xj=None, ni_normals=None, nj_normals=None, keep_lists=False):
        """Targets xi (n,3), sources xj (m,3) (default: the same points, which
        then share one tree).  Each array may also be a sequence of its three
        columns (x, y, z), the engine's own layout, which is then used
        without a copy when contiguous float32."""
        same = xj is None or xj is xi
        if xj is None:
            xj = xi

        def columns(a):
            if a is None:
                return [None] * 3
            if isinstance(a, (tuple, list)):
                if len(a) != 3:
                    raise ValueError("a sequence of columns needs 3 entries (x, y, z)")
                return list(a)
            return [a[:, k] for k in range(3)]

        ci, cj = columns(xi), columns(xj)
        dev = _is_torch(ci[0]) and ci[0].is_cuda
        self._device = dev
        cols = [_Arr(c, dev) for c in ci]
        cols += [_Arr(c, dev) for c in columns(ni_normals)]
        # the same column buffers for both sides: the engine then builds one tree
        cols += cols[:3] if same else [_Arr(c, dev) for c in cj]
        cols += [_Arr(c, dev) for c in columns(nj_normals)]
        self._cols = cols
        self.ni, self.nj = int(ci[0].shape[0]), int(cj[0].shape[0])
        flags = (1 if dev else 0) | (2 if keep_lists else 0)
        _check(lib().fmmx_set_points(self.h, self.ni, *[c.ptr for c in cols[:6]], self.nj,
                                     *[c.ptr for c in cols[6:]], flags))
        return self

    # -- matvec ---------------------------------------------------------
    def evaluate(self, q, neq=G, gradient=True):
        """Potential v (ni,) and, for neq=G or YUKAWA, gradient g (ni,3), caller order.
        HELMHOLTZ: complex charges in, complex potentials out (no gradient)."""
        dev = self._device
        if neq == HELMHOLTZ:
            return self._evaluate_complex(q), None
        qa = _Arr(q, dev)
        want_g = gradient and neq in (G, YUKAWA)
        if dev:
            import torch
            v = torch.empty(self.ni, dtype=torch.float64, device=q.device)
            g = torch.empty((3, self.ni), dtype=torch.float64, device=q.device) if want_g else None
            ptr = lambda t: t.data_ptr()  # noqa: E731
        else:
            v = np.empty(self.ni)
            g = np.empty((3, self.ni)) if want_g else None
            ptr = lambda t: t.ctypes.data  # noqa: E731
        gp = [ptr(g[k]) for k in range(3)] if want_g else [None] * 3
        _check(lib().fmmx_evaluate(self.h, neq, qa.ptr, ptr(v), *gp, 1 if dev else 0))
        if want_g:
            return v, g.T
        return v, None

    # -- adjoint ----------------------------------------------------------
    def _out(self, shape, like_dev):
        if like_dev:
            import torch
            return torch.empty(shape, dtype=torch.float64, device="cuda"), (lambda t: t.data_ptr())
        return np.empty(shape), (lambda t: t.ctypes.data)

    def transpose(self, w, neq=G, gradient=False):
        """u = K^T w (one value per source, caller order) for strengths w on
        the targets; for G and YUKAWA with gradient also the gradient of the
        transposed field at the sources (nj, 3).  HELMHOLTZ: complex w, u."""
        dev = self._device
        if neq == HELMHOLTZ:
            wp = self._pairs(w, dev)
            u, ptr = self._out((self.nj, 2), dev)
            _check(lib().fmmx_evaluate_transpose(self.h, neq, wp.data_ptr() if dev else wp.ctypes.data, ptr(u),
                                                 None, None, None, 1 if dev else 0))
            if dev:
                import torch
                return torch.view_as_complex(u), None
            return u[:, 0] + 1j * u[:, 1], None
        wa = _Arr(w, dev)
        want_g = gradient and neq in (G, YUKAWA)
        u, ptr = self._out(self.nj, dev)
        g, _ = self._out((3, self.nj), dev) if want_g else (None, None)
        gp = [ptr(g[k]) for k in range(3)] if want_g else [None] * 3
        _check(lib().fmmx_evaluate_transpose(self.h, neq, wa.ptr, ptr(u), *gp, 1 if dev else 0))
        return u, (g.T if want_g else None)

    def gradients(self, q, w, neq=G):
        """Gradients of L = w . (K q) for G (or YUKAWA): dL/dq (nj,),
        dL/dx_i (ni, 3) and dL/dx_j (nj, 3), caller order."""
        dev = self._device
        qa, wa = _Arr(q, dev), _Arr(w, dev)
        dq, ptr = self._out(self.nj, dev)
        dxi, _ = self._out((3, self.ni), dev)
        dxj, _ = self._out((3, self.nj), dev)
        _check(lib().fmmx_gradients(self.h, neq, qa.ptr, wa.ptr, ptr(dq), ptr(dxi), ptr(dxj), 1 if dev else 0))
        return dq, dxi.T, dxj.T

    def energy(self, q, neq=G):
        """E = 1/2 q . K q on one point set (targets = sources) and its
        gradients dE/dq (n,) and dE/dx (n, 3)."""
        dev = self._device
        qa = _Arr(q, dev)
        e = np.zeros(1)
        dq, ptr = self._out(self.ni, dev)
        dx, _ = self._out((3, self.ni), dev)
        if dev:
            import torch
            et = torch.zeros(1, dtype=torch.float64, device="cuda")
            _check(lib().fmmx_energy(self.h, neq, qa.ptr, et.data_ptr(), ptr(dq), ptr(dx), 1))
            return float(et.item()), dq, dx.T
        _check(lib().fmmx_energy(self.h, neq, qa.ptr, e.ctypes.data, ptr(dq), ptr(dx), 0))
        return float(e[0]), dq, dx.T

    def evaluate_phase(self, phase, q=None, neq=G, lev=0, gradient=True):
        """One phase of the matvec (P2P=0, P2M=1, M2M=2, M2L=3, L2L=4,
        L2P=5; fmmx_evaluate_phase).  P2P and L2P return (v, g); the others
        None."""
        dev = self._device
        qa = _Arr(q, dev) if q is not None else None
        writes = phase in (0, 5)
        want_g = writes and gradient and neq in (G, YUKAWA)
        v, ptr = self._out(self.ni, dev) if writes else (None, None)
        g, _ = self._out((3, self.ni), dev) if want_g else (None, None)
        gp = [ptr(g[k]) for k in range(3)] if want_g else [None] * 3
        _check(lib().fmmx_evaluate_phase(self.h, phase, neq, lev, qa.ptr if qa else None, ptr(v) if writes else None,
                                         *gp, 1 if dev else 0))
        return (v, g.T if want_g else None) if writes else None

    @staticmethod
    def _pairs(q, dev):
        """complex charges -> contiguous (re, im) float32 pairs"""
        if dev:
            import torch
            return torch.view_as_real(q.to(torch.complex64)).contiguous()
        q = np.asarray(q)
        return np.ascontiguousarray(np.stack([q.real, q.imag], -1).astype(np.float32))

    def _evaluate_complex(self, q):
        dev = self._device
        qp = self._pairs(q, dev)
        if dev:
            import torch
            v = torch.empty((self.ni, 2), dtype=torch.float64, device=q.device)
            _check(lib().fmmx_evaluate(self.h, HELMHOLTZ, qp.data_ptr(), v.data_ptr(), None, None, None, 1))
            return torch.view_as_complex(v)
        v = np.empty((self.ni, 2))
        _check(lib().fmmx_evaluate(self.h, HELMHOLTZ, qp.ctypes.data, v.ctypes.data, None, None, None, 0))
        return v[:, 0] + 1j * v[:, 1]

    def direct(self, xi, xj, q, neq=G, ni_normals=None, nj_normals=None):
        """GPU direct sum (FP64 accumulation) in this handle's convention.
        HELMHOLTZ: complex charges and potentials."""
        dev = _is_torch(xi) and xi.is_cuda
        if neq == HELMHOLTZ:
            a = [_Arr(xi[:, k], dev) for k in range(3)]
            b = [_Arr(xj[:, k], dev) for k in range(3)]
            qp = self._pairs(q, dev)
            n = int(xi.shape[0])
            if dev:
                import torch
                v = torch.zeros((n, 2), dtype=torch.float64, device=xi.device)
                _check(lib().fmmx_direct(self.h, neq, n, *[c.ptr for c in a], None, None, None, int(xj.shape[0]),
                                         *[c.ptr for c in b], None, None, None, qp.data_ptr(), v.data_ptr(), None, None,
                                         None, 1))
                return torch.view_as_complex(v), None
            v = np.zeros((n, 2))
            _check(lib().fmmx_direct(self.h, neq, n, *[c.ptr for c in a], None, None, None, int(xj.shape[0]),
                                     *[c.ptr for c in b], None, None, None, qp.ctypes.data, v.ctypes.data, None, None,
                                     None, 0))
            return v[:, 0] + 1j * v[:, 1], None
        a = [_Arr(xi[:, k], dev) for k in range(3)]
        a += [_Arr(None if ni_normals is None else ni_normals[:, k], dev) for k in range(3)]
        b = [_Arr(xj[:, k], dev) for k in range(3)]
        b += [_Arr(None if nj_normals is None else nj_normals[:, k], dev) for k in range(3)]
        qa = _Arr(q, dev)
        n = int(xi.shape[0])
        if dev:
            import torch
            v = torch.zeros(n, dtype=torch.float64, device=xi.device)
            g = torch.zeros((3, n), dtype=torch.float64, device=xi.device)
            ptr = lambda t: t.data_ptr()  # noqa: E731
        else:
            v = np.zeros(n)
            g = np.zeros((3, n))
            ptr = lambda t: t.ctypes.data  # noqa: E731
        _check(lib().fmmx_direct(self.h, neq, n, *[c.ptr for c in a], int(xj.shape[0]), *[c.ptr for c in b], qa.ptr,
                                 ptr(v), ptr(g[0]), ptr(g[1]), ptr(g[2]), 1 if dev else 0))
        return v, (g.T if neq in (G, YUKAWA) else None)

    def synchronize(self):
        _check(lib().fmmx_synchronize(self.h))

    def set_timing(self, per_phase=True):
        _check(lib().fmmx_set_timing(self.h, int(per_phase)))

    def stats(self):
        s = Stats()
        _check(lib().fmmx_get_stats(self.h, C.byref(s)))
        d = s.as_dict()
        # whether ms_m2l_leaf timed one launch over every level
        d["m2l_all_levels"] = lib().fmmx_get_count(self.h, 2, 0, 0) == 1
        return d

    # -- introspection ----------------------------------------------------
    @property
    def lmax(self):
        return lib().fmmx_get_count(self.h, 0, 0, 0)

    def domain(self):
        """(lmax, (xmin, ymin, zmin, rd)) of the tree in use."""
        lm = C.c_int(0)
        b = np.zeros(4, np.float32)
        _check(lib().fmmx_get_domain(self.h, C.byref(lm), b.ctypes.data))
        return lm.value, tuple(float(v) for v in b)

    def occupancy(self):
        """(boxes, sumsq) of levels 0..8 of the points the last set_points
        chose its depth from: boxes holding a point and the sums of squared
        points per box (zeros when lmax was given or under REF)."""
        b = np.zeros(9, np.int64)
        q = np.zeros(9, np.float64)
        _check(lib().fmmx_get_occupancy(self.h, b.ctypes.data, q.ctypes.data))
        return b, q

    def nboxes(self, side, lev):
        return lib().fmmx_get_count(self.h, 1, 0 if side == "i" else 1, lev)

    def perm(self, side):
        n = self.ni if side == "i" else self.nj
        a = np.zeros(n, np.int32)
        _check(lib().fmmx_get_perm(self.h, 0 if side == "i" else 1, a.ctypes.data))
        return a

    def boxes(self, side, lev):
        n = self.nboxes(side, lev)
        k, f, l = (np.zeros(n, np.int32) for _ in range(3))
        _check(lib().fmmx_get_boxes(self.h, 0 if side == "i" else 1, lev, k.ctypes.data, f.ctypes.data,
                                    l.ctypes.data))
        return k, f, l

    def interaction_list(self, lev, kind):
        n = self.nboxes("i", lev)
        nij = np.zeros(n, np.int32)
        neij = np.zeros(n * NEBM, np.int32)
        _check(lib().fmmx_get_list(self.h, lev, kind, nij.ctypes.data, neij.ctypes.data))
        neij = neij.reshape(n, NEBM)
        flat = np.concatenate([neij[i, :nij[i]] for i in range(n)]) if n else np.zeros(0, np.int32)
        return nij, flat

    def multipoles(self):
        nr = self.p * (self.p + 1) // 2
        rows = sum(self.nboxes("j", l) for l in range(2, self.lmax + 1))
        a = np.zeros(rows * nr * 2)
        _check(lib().fmmx_get_multipoles(self.h, a.ctypes.data))
        return a.view(np.complex128).reshape(rows, nr)

    def locals(self, lev):
        nr = self.p * (self.p + 1) // 2
        a = np.zeros(self.nboxes("i", lev) * nr * 2)
        _check(lib().fmmx_get_locals(self.h, lev, a.ctypes.data))
        return a.view(np.complex128).reshape(-1, nr)


def matvec(xi, q, xj=None, neq=G, **kw):
    """One-shot: tree set-up and one evaluation."""
    f = FMM(**kw)
    try:
        f.set_points(xi, xj)
        return f.evaluate(q, neq=neq)
    finally:
        f.close()
