"""Reverse-mode rule of the solve x = A^-1 b on the GPU.

The reference has no AD rules; this is the ``rrule`` a ChainRulesCore method
on ``amg_solve!`` would carry, written as ``torch.autograd.Function``s over the
C-ABI adjoint entry points (``draugr_hip_assemble_7pt``,
``draugr_hip_matrix_grad``, ``draugr_hip_grad_logk_7pt``).

Forward: the operator's values are written into the hierarchy (partial
resetup, src/resetup.jl:20-45: P, maps and patterns fixed) and x is found by
the FP64 PCG with one AMG cycle as preconditioner.

Backward, for an upstream gradient g = dL/dx:
  * lambda = A^-T g, solved with the same hierarchy (A is symmetric);
  * dL/db = lambda;
  * dL/dA_ij = -lambda_i x_j on the pattern of A;
  * for the 7-point operator of cell permeabilities k = exp(log k):
    dL/dlog k_c = sum over faces f of c of dL/dT_f * k_c dT_f/dk_c plus the
    Dirichlet half-cell terms, with dL/dT_f = -(l_a - l_b)(x_a - x_b) and
    k_a dT/dk_a = 2 k_a k_b^2 / (k_a + k_b)^2 (T = 2 k_a k_b / (k_a + k_b)).

The gradient is that of the exact solve; PCG stops at ``tol`` in both
directions, so the gradient carries an error of the order of ``tol`` times the
condition number.
"""
from __future__ import annotations

from typing import Optional

from . import _lib
from ._lib import DraugrError, check
from .config import AMGConfig
from .core import AMGHierarchy, _dptr, _torch_stream, amg_resetup_, pcg_solve_


def finest_nnz(h: AMGHierarchy) -> int:
    if h.nlevels > 0:
        return h.level_info(0)["nnz"]
    return len(h.coarsest()["nz"])


def assemble_7pt_(h: AMGHierarchy, m: int, k, nzval=None):
    """Values of the 7-point operator of permeabilities ``k`` (m^3 cells) in the
    layout of the hierarchy's finest operator, computed on the device."""
    import torch
    nnz = finest_nnz(h)
    if nzval is None:
        nzval = torch.empty(nnz, dtype=torch.float64, device=k.device)
    check(_lib.lib().draugr_hip_assemble_7pt(h.handle, m, _dptr(k, m ** 3, "k"), _dptr(nzval, nnz, "nzval"),
                                             _torch_stream()), "draugr_hip_assemble_7pt")
    return nzval


def matrix_grad(h: AMGHierarchy, x, lam, out=None):
    """dL/dA on the finest operator's pattern: -lam_i x_j per stored entry."""
    import torch
    nnz = finest_nnz(h)
    if out is None:
        out = torch.empty(nnz, dtype=torch.float64, device=x.device)
    check(_lib.lib().draugr_hip_matrix_grad(h.handle, _dptr(x, h.n, "x"), _dptr(lam, h.n, "lambda"),
                                            _dptr(out, nnz, "grad"), _torch_stream()), "draugr_hip_matrix_grad")
    return out


def grad_logk_7pt(m: int, k, x, lam, out=None):
    """dL/dlog k of the 7-point operator from the solution x and adjoint lam."""
    import torch
    n = m ** 3
    if out is None:
        out = torch.empty(n, dtype=torch.float64, device=k.device)
    check(_lib.lib().draugr_hip_grad_logk_7pt(m, _dptr(k, n, "k"), _dptr(x, n, "x"), _dptr(lam, n, "lambda"),
                                              _dptr(out, n, "grad"), _torch_stream()), "draugr_hip_grad_logk_7pt")
    return out


class _State:
    """Which operator values the hierarchy currently holds (so that a backward
    after another forward on the same hierarchy resets it up again)."""

    def __init__(self):
        self.token = None


def _state(h: AMGHierarchy) -> _State:
    s = getattr(h, "_adjoint_state", None)
    if s is None:
        s = _State()
        h._adjoint_state = s
    return s


def _pcg(h, b, config, tol, maxiter):
    import torch
    x = torch.zeros_like(b)
    _, it, rel = pcg_solve_(x, b, h, config, tol=tol, maxiter=maxiter)
    if not rel < tol:
        raise DraugrError("PCG did not reach tol=%g in %d iterations (rel %g)" % (tol, maxiter, rel))
    return x, it


def _make_functions():
    import torch

    class _Solve7pt(torch.autograd.Function):
        @staticmethod
        def forward(ctx, logk, b, h, m, config, tol, maxiter):
            k = torch.exp(logk.detach()).contiguous()
            nz = assemble_7pt_(h, m, k)
            amg_resetup_(h, nz, config)
            st = _state(h)
            st.token = object()
            x, it = _pcg(h, b.detach().contiguous(), config, tol, maxiter)
            ctx.save_for_backward(k, x, nz)
            ctx.h, ctx.m, ctx.config, ctx.tol, ctx.maxiter, ctx.token = h, m, config, tol, maxiter, st.token
            ctx.iterations = it
            return x

        @staticmethod
        def backward(ctx, gx):
            k, x, nz = ctx.saved_tensors
            h = ctx.h
            st = _state(h)
            if st.token is not ctx.token:  # another forward changed the operator
                amg_resetup_(h, nz, ctx.config)
                st.token = ctx.token
            lam, _ = _pcg(h, gx.contiguous(), ctx.config, ctx.tol, ctx.maxiter)
            glogk = grad_logk_7pt(ctx.m, k, x, lam) if ctx.needs_input_grad[0] else None
            gb = lam if ctx.needs_input_grad[1] else None
            return glogk, gb, None, None, None, None, None

    class _SparseSolve(torch.autograd.Function):
        @staticmethod
        def forward(ctx, nzval, b, h, config, tol, maxiter):
            nz = nzval.detach().contiguous()
            amg_resetup_(h, nz, config)
            st = _state(h)
            st.token = object()
            x, it = _pcg(h, b.detach().contiguous(), config, tol, maxiter)
            ctx.save_for_backward(x, nz)
            ctx.h, ctx.config, ctx.tol, ctx.maxiter, ctx.token = h, config, tol, maxiter, st.token
            ctx.iterations = it
            return x

        @staticmethod
        def backward(ctx, gx):
            x, nz = ctx.saved_tensors
            h = ctx.h
            st = _state(h)
            if st.token is not ctx.token:
                amg_resetup_(h, nz, ctx.config)
                st.token = ctx.token
            lam, _ = _pcg(h, gx.contiguous(), ctx.config, ctx.tol, ctx.maxiter)
            gnz = matrix_grad(h, x, lam) if ctx.needs_input_grad[0] else None
            gb = lam if ctx.needs_input_grad[1] else None
            return gnz, gb, None, None, None, None

    return _Solve7pt, _SparseSolve


_FUNCS = None


def _funcs():
    global _FUNCS
    if _FUNCS is None:
        _FUNCS = _make_functions()
    return _FUNCS


def solve_7pt(logk, b, h: AMGHierarchy, m: int, config: Optional[AMGConfig] = None, tol: float = 1e-10,
              maxiter: int = 1000):
    """x = A(exp(logk))^-1 b for the 7-point operator on m^3 cells, differentiable
    in ``logk`` and ``b`` (float64 CUDA tensors).  ``h`` is a hierarchy set up
    on a 7-point operator of the same grid; its values are replaced."""
    return _funcs()[0].apply(logk, b, h, m, config or h.config, tol, maxiter)


def sparse_solve(nzval, b, h: AMGHierarchy, config: Optional[AMGConfig] = None, tol: float = 1e-10,
                 maxiter: int = 1000):
    """x = A^-1 b with A the hierarchy's finest pattern carrying ``nzval``,
    differentiable in ``nzval`` and ``b``.  The pattern must be symmetric
    (values symmetric) for the adjoint solve with the same hierarchy."""
    return _funcs()[1].apply(nzval, b, h, config or h.config, tol, maxiter)
