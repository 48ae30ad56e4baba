"""Hand-written adjoints of the GPU kernels behind torch.autograd.Function (FP64).

No reference counterpart: the reference has no derivative capability on this path.

* ``ReducedSolve``  x = A(theta)^{-1} b for a batch of affine systems.  Backward solves
  the adjoint systems A(theta)^T lambda = dL/dx in one batched launch and returns
  dL/dtheta_q = -lambda^T A_q x and dL/db = lambda.
* ``BwdTransFn``    u = B u_hat; backward is the Euclidean transpose B^T (an IProduct
  with unit weights).
* ``L2Misfit``      0.5 * ||B u_hat - f||^2 in the quadrature-weighted L2 norm
  (sum_q w_q J_q r_q^2); its gradient is IProductWRTBase(B u_hat - f), evaluated by the
  IProduct kernel itself.
"""
import torch

from . import rom as _rom


class ReducedSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, b, batch, A, method=_rom.LU_FP64):
        x = batch.solve(A, theta.detach().contiguous(), b.detach().contiguous(), method=method)
        ctx.batch, ctx.A, ctx.method = batch, A, method
        ctx.save_for_backward(theta, x)
        return x

    @staticmethod
    def backward(ctx, gx):
        theta, x = ctx.saved_tensors
        lam = ctx.batch.solve(ctx.A, theta.detach().contiguous(), gx.contiguous(),
                              method=ctx.method, transpose=True)
        gtheta = ctx.batch.param_grad(ctx.A, lam, x)
        return gtheta, lam, None, None, None


def reduced_solve(batch, A, theta, b, method=_rom.LU_FP64):
    return ReducedSolve.apply(theta, b, batch, A, method)


class BwdTransFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u_hat, op):
        ctx.op = op
        return op.apply(u_hat.detach().contiguous())

    @staticmethod
    def backward(ctx, g):
        return ctx.op.apply_transpose(g.contiguous()), None


class L2Misfit(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u_hat, f, bwd, ipr, jw):
        r = bwd.apply(u_hat.detach().contiguous()) - f
        ctx.ipr = ipr
        ctx.save_for_backward(r)
        return 0.5 * torch.dot(jw * r, r)

    @staticmethod
    def backward(ctx, g):
        (r,) = ctx.saved_tensors
        return g * ctx.ipr.apply(r), None, None, None, None
