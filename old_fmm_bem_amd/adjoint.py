"""Adjoints of the FMM matvec and a torch.autograd wrapper.

The matvec phi = K q of kernel/G_p2p.cxx (K_ij = G(x_i - x_j), G(r) =
1/(4 pi |r|)) is symmetric when targets and sources are the same points, and
its derivative with respect to the points is the field gradient the forward
pass already returns (kernel/G_l2p.cxx:68-70 writes it as gxi, gyi, gzi).
For L = sum_i w_i phi_i:

    dL/dq_k = (K^T w)_k = (K w)_k
    dL/dx_k = w_k grad phi_q(x_k) + q_k grad phi_w(x_k)

(G is even, so the gradient of G(x_i - x_k) with respect to x_k is minus the
gradient with respect to x_i).  The backward pass therefore runs one more
matvec with strengths w on the tree of the forward pass; expansions are
recomputed, not stored.  For the double-layer kernels on one point set with
one set of normals, K_Gni^T = -K_Gnj (kernel/Gni_p2p.cxx vs Gnj_p2p.cxx:
(x_i - x_j).n_i = -(x_j - x_i).n_i).
"""
from __future__ import annotations

import numpy as np

from . import FMM, G, GNI, GNJ


def adjoint(f: FMM, w, neq=G):
    """K^T w for a handle whose targets and sources are the same points
    (set_points(x) or set_points(x, x, ...)); for Gni/Gnj the handle must
    hold the same normals on both sides."""
    if neq == G:
        v, _ = f.evaluate(w, neq=G, gradient=False)
        return v
    v, _ = f.evaluate(w, neq=GNJ if neq == GNI else GNI)
    return -v


def energy_and_gradients(f: FMM, q):
    """E = 1/2 q.K q, dE/dq = K q and dE/dx_k = q_k grad phi(x_k) for the
    single-layer kernel on the handle's point set (targets = sources)."""
    v, g = f.evaluate(q, neq=G)
    if isinstance(v, np.ndarray):
        qd = np.asarray(q, np.float64)
        return 0.5 * float(qd @ v), v, qd[:, None] * g
    qd = q.to(v.dtype)
    return 0.5 * float((qd * v).sum()), v, qd[:, None] * g


def potential(x, q, p=10, lmax=0, precision="mixed", convention="exact", box=None, device=0):
    """Differentiable phi = K q (float64) on the points x (targets = sources),
    torch CUDA tensors x (n, 3) and q (n,).  Gradients flow to q and x."""
    import torch

    class _Potential(torch.autograd.Function):
        @staticmethod
        def forward(ctx, xx, qq):
            f = FMM(p=p, lmax=lmax, precision=precision, convention=convention, box=box, device=device)
            f.set_points(xx.detach())
            v, g = f.evaluate(qq.detach().to(torch.float32))
            ctx.f = f
            ctx.save_for_backward(qq, g)
            return v

        @staticmethod
        def backward(ctx, w):
            qq, g_q = ctx.saved_tensors
            f = ctx.f
            v_w, g_w = f.evaluate(w.detach().to(torch.float32).contiguous())
            f.close()
            ctx.f = None
            gx = w[:, None] * g_q + qq.to(torch.float64)[:, None] * g_w
            return gx.to(x.dtype), v_w.to(qq.dtype)

    return _Potential.apply(x, q)
