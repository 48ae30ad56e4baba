"""Hand-written adjoints of the GPU kernels behind torch.autograd.Function (FP64).

No reference counterpart: the reference has no derivative capability on this path.

* ``ReducedSolve``  x = A(theta)^{-1} b for a batch of affine systems.  Backward solves
  the adjoint systems A(theta)^T lambda = dL/dx in one batched launch and returns
  dL/dtheta_q = -lambda^T A_q x and dL/db = lambda.
* ``BwdTransFn``    u = B u_hat; backward is the Euclidean transpose B^T (an IProduct
  with unit weights).
* ``L2Misfit``      0.5 * ||B u_hat - f||^2 in the quadrature-weighted L2 norm
  (sum_q w_q J_q r_q^2); its gradient in u_hat is IProductWRTBase(B u_hat - f), evaluated
  by the IProduct kernel itself, and its gradient in f is -diag(w J)(B u_hat - f).
* ``HelmSolveFn``   u = A(lambda)^{-1} rhs, the assembled Helmholtz solve with homogeneous
  Dirichlet modes (A = K + lambda M on the free modes, symmetric).  Backward is one more
  solve, w = A^{-1} dL/du, and returns dL/drhs = w and dL/dlambda = -w^T M u, with M u
  assembled from the element mass operator.
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
        r = bwd.apply(u_hat.detach().contiguous()) - f.detach()
        ctx.ipr = ipr
        ctx.save_for_backward(r, jw)
        return 0.5 * torch.dot(jw * r, r)

    @staticmethod
    def backward(ctx, g):
        r, jw = ctx.saved_tensors
        gu = g * ctx.ipr.apply(r) if ctx.needs_input_grad[0] else None
        gf = -g * jw * r if ctx.needs_input_grad[1] else None
        return gu, gf, None, None, None


def _set_lambda(op, lam):
    for o in (op if isinstance(op, (list, tuple)) else [op]):
        o.set_lambda(lam)


class HelmSolveFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rhs, lam, asm, op, mass, tol):
        _set_lambda(op, float(lam))
        x, _, _ = asm.helmsolve(op, rhs.detach().contiguous(), tol=tol)
        ctx.asm, ctx.op, ctx.mass, ctx.tol, ctx.lam = asm, op, mass, tol, float(lam)
        ctx.lam_like = lam if torch.is_tensor(lam) else None
        ctx.save_for_backward(x)
        return x

    @staticmethod
    def backward(ctx, g):
        (x,) = ctx.saved_tensors
        _set_lambda(ctx.op, ctx.lam)
        w, _, _ = ctx.asm.helmsolve(ctx.op, g.contiguous(), tol=ctx.tol)
        glam = None
        if ctx.needs_input_grad[1]:
            if ctx.mass is None:
                raise ValueError("HelmSolveFn: a mass operator is needed for d/dlambda")
            mu = ctx.asm.assemble(ctx.asm.apply_local(ctx.mass, ctx.asm.global_to_local(x)))
            ref = ctx.lam_like
            glam = (-torch.dot(w, mu)).to(device=ref.device, dtype=ref.dtype).reshape(ref.shape)
        return w, glam, None, None, None, None


def helm_solve(asm, op, rhs, lam, mass=None, tol=1e-12):
    """u = A(lambda)^{-1} rhs through the assembly `asm` and Helmholtz operator `op`
    (differentiable in rhs and, given the element mass operator `mass`, in lambda, a
    0-d tensor).  For a variable-order assembly, `op` and `mass` are lists with one
    operator per run of equal-order elements, as Assembly.helmsolve takes them.  The
    operators' lambda is set to `lam` (and left there)."""
    return HelmSolveFn.apply(rhs, lam, asm, op, mass, tol)
