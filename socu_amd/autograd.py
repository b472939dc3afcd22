"""Differentiable block-tridiagonal solve (no counterpart in the reference,
whose launches are not differentiable).

``block_tridiag_solve(L, E, b)`` returns x = Psi^{-1} b for the matrix the
reference's solver factors (``block_tridiag_solver.py:739-1727``): diagonal
blocks built from the lower triangle of each ``L[i]`` (the upper triangle is
ignored), coupling block ``E[i]`` at block position (i+1, i) and its
transpose at (i, i+1).  The fill-in region of ``E`` (entries from N-1 on) is
workspace: the forward pass factors a copy of L and of the coupling blocks
(the factor overwrites the fill-in region), so the inputs are never modified
and the fill-in entries carry no gradient.

Backward reuses the forward factor for the transpose solve
bbar = Psi^{-1} xbar (Psi is symmetric) and then one HIP kernel scatters
-bbar x^T onto the parameters (``socu_adjoint_scatter``).
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from .block_tridiag_solver import _batched


def _code(t):
    if t.dtype == torch.float64:
        return _lib.DTYPE_F64
    if t.dtype == torch.float32:
        return _lib.DTYPE_F32
    raise TypeError(f"block_tridiag_solve supports float32 and float64, got {t.dtype}")


def _stream(dev):
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


class BlockTridiagSolve(torch.autograd.Function):
    """x = Psi(L, E)^{-1} b with gradients for L, E and b."""

    @staticmethod
    def forward(ctx, L, E, b):
        for name, t in (("L", L), ("E", E), ("b", b)):
            if not isinstance(t, torch.Tensor):
                raise TypeError(f"{name} must be a torch.Tensor")
            if t.device.type != "cuda":
                raise ValueError(f"{name} is on {t.device}: socu_amd has no CPU implementation")
        if not (L.dtype == E.dtype == b.dtype):
            raise TypeError("L, E and b must share one dtype")
        if not (L.device == E.device == b.device):
            raise ValueError("L, E and b must be on one device")
        shape_b = b.shape
        L4, E4, b4, B, N, n, m = _batched(L.detach(), E.detach(), b.detach())
        code = _code(L)
        lib = _lib.load()
        # the factor goes to fresh buffers, read straight from the inputs
        # (out of place: no copies of L and E), which stay as given
        Ls = L4.contiguous()
        Es = E4.contiguous()
        Lw = torch.empty_like(Ls)
        Ew = torch.empty_like(Es)
        x = b4.contiguous().clone()
        with torch.cuda.device(L.device):
            _lib.check(lib.socu_cholesky_factor_and_solve_from(Ls.data_ptr(), Es.data_ptr(), Lw.data_ptr(),
                                                               Ew.data_ptr(), x.data_ptr(), None, B, N, n, m, code,
                                                               _stream(L.device)), "block_tridiag_solve")
        ctx.save_for_backward(Lw, Ew, x)
        ctx.dims = (B, N, n, m, code, L.shape, E.shape, shape_b)
        return x.reshape(shape_b)

    @staticmethod
    def backward(ctx, xbar):
        Lw, Ew, x = ctx.saved_tensors
        B, N, n, m, code, shape_L, shape_E, shape_b = ctx.dims
        lib = _lib.load()
        dev = x.device
        bbar = xbar.reshape(B, N, n, m).to(dtype=x.dtype).contiguous().clone()
        with torch.cuda.device(dev):
            # Lw is this library's factor: its diagonal-block inverses are in
            # place, so the transpose solve needs no inverse launch
            _lib.check(lib.socu_cholesky_solve_with_inverses(Lw.data_ptr(), Ew.data_ptr(), bbar.data_ptr(), None, B, N,
                                                             n, m, code, _stream(dev)),
                       "block_tridiag_solve backward (solve)")
            dL = dE = None
            if ctx.needs_input_grad[0] or ctx.needs_input_grad[1]:
                dL = torch.empty_like(Lw)
                dE = torch.empty_like(Ew)
                _lib.check(lib.socu_adjoint_scatter(bbar.data_ptr(), x.data_ptr(), dL.data_ptr(), dE.data_ptr(),
                                                    None, B, N, n, m, code, _stream(dev)),
                           "block_tridiag_solve backward (scatter)")
                dL = dL.reshape(shape_L) if ctx.needs_input_grad[0] else None
                dE = dE.reshape(shape_E) if ctx.needs_input_grad[1] else None
        db = bbar.reshape(shape_b) if ctx.needs_input_grad[2] else None
        return dL, dE, db


def block_tridiag_solve(L: torch.Tensor, E: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """Differentiable x = Psi^{-1} b; shapes as the solver's launches:
    L (N,n,n) or (B,N,n,n), E (n_E,n,n) or (B,n_E,n,n), b (N,n,m) or
    (B,N,n,m).  Inputs are not modified."""
    return BlockTridiagSolve.apply(L, E, b)
