"""torch.autograd.Function over the HIP sweep and its adjoint.

Used only to check the adjoint: forward = T steps of the stencil on a single
grid (halo fixed), backward = the temporally blocked adjoint sweep seeded
with the incoming gradient.  The adjoint entry point differentiates
L = 1/2 ||u_T - target||^2 over the interior, so a vector-Jacobian product
with an arbitrary g is obtained with target = u_T - g (lambda_T = g on the
interior).  Gradients flow to the initial interior and to the coefficients
(gradient2d: {eps, scale}); the halo is a fixed boundary condition.
"""
import numpy as np
import torch

from . import AdjointWorkspace, adjoint_device, forward_device, info


class StencilSweep(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0, coefs, name, n, T, divisor=-1.0, bt=0):
        c = coefs.detach().cpu().double().numpy()
        grid = torch.stack([u0.detach(), u0.detach()]).contiguous()
        forward_device(name, grid, n, T, coefs=c, divisor=divisor, bt=bt)
        uT = grid[T % 2].clone()
        ctx.save_for_backward(u0, coefs, uT)
        ctx.cfg = (name, list(n), T, divisor, bt)
        return uT

    @staticmethod
    def backward(ctx, g):
        u0, coefs, uT = ctx.saved_tensors
        name, n, T, divisor, bt = ctx.cfg
        d = info(name)
        dt = np.float32 if u0.dtype == torch.float32 else np.float64
        ws = AdjointWorkspace(name, n, T, dt, bt=bt)
        target = (uT - g).contiguous()
        grad = torch.zeros_like(u0)
        gc = torch.zeros(d["ntaps"] if d["linear"] else 2, dtype=torch.float64, device=u0.device)
        loss = torch.zeros(1, dtype=torch.float64, device=u0.device)
        c = coefs.detach().cpu().double().numpy()
        adjoint_device(name, u0.detach().contiguous(), target, grad, gc, loss, n, T, ws, coefs=c,
                       divisor=divisor, bt=bt)
        return grad, gc.to(coefs.dtype), None, None, None, None, None


def sweep(u0, coefs, name, n, T, divisor=-1.0, bt=0):
    return StencilSweep.apply(u0, coefs, name, n, T, divisor, bt)
