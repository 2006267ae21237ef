"""Adjoints of the FMM matvec and a torch.autograd wrapper, over the C ABI
(fmmx_evaluate_transpose, fmmx_gradients, fmmx_energy in include/fmmx.h).

For v = K q (kernel/G_p2p.cxx: K_ij = G(x_i - x_j), G(r) = 1/(4 pi |r|))
and L = sum_i w_i v_i:

    dL/dq   = K^T w                       (the transposed matvec)
    dL/dx_i = w_i grad phi_q(x_i)          (kernel/G_l2p.cxx:68-70)
    dL/dx_j = q_j grad psi_w(x_j),  psi_w = K^T w

The engine computes all three on the device: one forward evaluation with
the field gradient and one transposed evaluation on a second tree whose
targets are the sources (expansions recomputed, not stored).  The transpose
of the target-normal kernel Gni (kernel/Gni_p2p.cxx) is minus the
source-normal kernel Gnj (kernel/Gnj_p2p.cxx) carrying the target normals.
"""
from __future__ import annotations

from . import FMM, G


def adjoint(f: FMM, w, neq=G):
    """K^T w: one value per source of the handle's tree (caller order)."""
    return f.transpose(w, neq=neq)[0]


def energy_and_gradients(f: FMM, q, neq=G):
    """E = 1/2 q.K q, dE/dq = K q and dE/dx_k = q_k grad phi(x_k) on the
    handle's point set (targets = sources)."""
    return f.energy(q, neq=neq)


def potential(x, q, xj=None, p=10, lmax=0, precision="mixed", convention="exact", box=None, device=0):
    """Differentiable phi = K q (float64) at the targets x from charges q at
    the sources xj (default: the targets themselves), torch CUDA tensors.
    Gradients flow to q, x and xj through fmmx_gradients."""
    import torch

    same = xj is None

    class _Potential(torch.autograd.Function):
        @staticmethod
        def forward(ctx, xx, qq, yy):
            f = FMM(p=p, lmax=lmax, precision=precision, convention=convention, box=box, device=device)
            f.set_points(xx.detach(), None if same else yy.detach())
            v, _ = f.evaluate(qq.detach().to(torch.float32), gradient=False)
            ctx.f = f
            ctx.save_for_backward(qq)
            return v

        @staticmethod
        def backward(ctx, w):
            (qq,) = ctx.saved_tensors
            f = ctx.f
            dq, dxi, dxj = f.gradients(qq.detach().to(torch.float32), w.detach().to(torch.float32).contiguous())
            f.close()
            ctx.f = None
            if same:
                return (dxi + dxj).to(x.dtype), dq.to(qq.dtype), None
            return dxi.to(x.dtype), dq.to(qq.dtype), dxj.to(xj.dtype)

    return _Potential.apply(x, q, x if same else xj)
