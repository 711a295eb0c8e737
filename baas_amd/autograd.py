"""Differentiable sparse solves (torch.autograd) on the HIP kernels.

* :class:`TriSolve` -- x = L^-1 b for a lower-triangular L on a schedule.
  Backward: lambda = L^-T g is the transposed (upper) sweep over the reversed
  DAG with the same execution plan (the reference's backward P2P machinery,
  merge/SpTRSV.cpp:601-660, reused as the adjoint), then
  dL_ij = -lambda_i x_j on L's pattern and db = lambda.
* :class:`ICCGSolve` -- x = A^-1 b by IC0-preconditioned CG (SPD A).
  Backward: lambda = A^-1 g by a second PCG solve on the same (symmetric) A,
  dA_ij = -lambda_i x_j on A's pattern and db = lambda.  The preconditioner
  drops out of the gradient because both solves are run to convergence.

Every stored nonzero is treated as an independent parameter.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib as L
from .api import PCG, Handle, Schedule, _ptr


class TriSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vals, b, sched: Schedule):
        x = sched.solve(vals, b)
        ctx.sched = sched
        ctx.save_for_backward(vals, x)
        return x

    @staticmethod
    def backward(ctx, g):
        vals, x = ctx.saved_tensors
        s = ctx.sched
        uv = s.transpose_values(vals)
        lam = s.solve(uv, g.contiguous(), uplo=1)
        dvals = s.grad_vals(lam, x) if ctx.needs_input_grad[0] else None
        return dvals, lam, None


def trisolve(vals, b, sched: Schedule):
    return TriSolve.apply(vals, b, sched)


class ICCGSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, a_vals, b, solver: PCG, tol: float, maxiter: int):
        # The PCG handle factorises and multiplies with the buffer it was
        # created on; any other tensor would be ignored by the forward pass
        # while still receiving the gradient, so it is refused.
        sa = solver.a_vals
        if (not isinstance(a_vals, torch.Tensor) or a_vals.dtype != sa.dtype or a_vals.numel() != sa.numel()
                or a_vals.device != sa.device or a_vals.data_ptr() != sa.data_ptr()):
            raise ValueError("iccg_solve: a_vals must be the tensor the PCG solver was created on "
                             "(same device buffer); create a PCG on this tensor instead")
        x, k, rel = solver.solve(b, tol=tol, maxiter=maxiter, refactor=True)
        ctx.solver, ctx.tol, ctx.maxiter = solver, tol, maxiter
        ctx.save_for_backward(a_vals, x)
        ctx.iterations = k
        return x

    @staticmethod
    def backward(ctx, g):
        a_vals, x = ctx.saved_tensors
        solver = ctx.solver
        lam, _, _ = solver.solve(g.contiguous(), tol=ctx.tol, maxiter=ctx.maxiter)
        da = None
        if ctx.needs_input_grad[0]:
            s = solver.s
            da = torch.empty_like(a_vals)
            L.call("baas_csr_outer_grad_f64", solver.h.ptr, s.n, _ptr(s.rowptr), _ptr(s.colidx), _ptr(lam),
                   _ptr(x), _ptr(da), 1)
        return da, lam, None, None, None


def iccg_solve(a_vals, b, solver: PCG, tol: float = 1e-12, maxiter: int = 10000):
    """x = A^-1 b with gradients for A's stored values and b.

    ``solver`` must have been created on ``a_vals`` (same device buffer): the
    forward pass refactors IC0 from it, so in-place updates are picked up.
    """
    return ICCGSolve.apply(a_vals, b, solver, tol, maxiter)


__all__ = ["TriSolve", "trisolve", "ICCGSolve", "iccg_solve", "Handle", "Schedule", "PCG", "C"]
