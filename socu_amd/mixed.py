"""Mixed-precision factor_and_solve with FP64 iterative refinement.

No counterpart in the reference (socu offers float32 or float64 end to end,
README.md:274-285); this is an extension on the same data layout:

* L (N,n,n) / (B,N,n,n) and E (n_E,n,n) / (B,n_E,n,n), float64, describe Psi
  exactly as for ``create_cholesky_factor_launch`` and are **not modified**;
* x (N,n,m) / (B,N,n,m), float64, holds b on entry and the solution on exit.

The launch owns an FP32 working copy of (L, E), factors it with the same ND
kernels, then refines.  ``precision``:

* ``"float32"`` (default): Schur updates on FP32 MFMA; the diagonal blocks are
  factored and inverted in FP64 and rounded to FP32;
* ``"bfloat16"``: as ``"float32"``, but the Schur-complement products (syrk
  and the fill-in gemm) run on BF16 MFMA with FP32 accumulation, their
  operands rounded from FP32 to BF16;
* ``"float32_f32_pivots"``: everything in FP32.

Refinement:
d = solve32(r), x += d, r = b - Psi x in FP64, until ||r||_F <= tol ||b||_F
per system or ``max_iter`` correction solves.  The convergence test runs on
the device; with ``use_cuda_graph`` the whole call is one hipGraph launch.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from .block_tridiag_solver import _batched, _check_mask, _check_tensors, _ptr, _resolve_device

__all__ = ["create_cholesky_mixed_factor_and_solve_launch"]

_PRECISIONS = {"float32": _lib.PREC_FP32, "float32_f32_pivots": _lib.PREC_FP32_F32PIV, "bfloat16": _lib.PREC_BF16}


class _MixedLaunch:
    def __init__(self, L, E, x, mask, B, N, n, m, device, stream, precision, tol, max_iter, use_graph):
        self._lib = _lib.load()
        self._keep = (L, E, x, mask)
        self._device = device
        self._stream = stream
        self._B = B
        self._h = ctypes.c_void_p()
        with torch.cuda.device(device):
            rc = self._lib.socu_mixed_create(ctypes.byref(self._h), _ptr(L), _ptr(E), _ptr(x), _ptr(mask), B, N, n, m,
                                             _PRECISIONS[precision], float(tol), int(max_iter), int(bool(use_graph)))
        _lib.check(rc, "cholesky_mixed_factor_and_solve")

    def _st(self):
        return self._stream if self._stream is not None else torch.cuda.current_stream(self._device)

    def __call__(self) -> None:
        with torch.cuda.device(self._device):
            rc = self._lib.socu_mixed_execute(self._h, ctypes.c_void_p(self._st().cuda_stream))
        _lib.check(rc, "cholesky_mixed_factor_and_solve")

    def stats(self):
        """(relative residual ||b - Psi x||_F / ||b||_F, correction solves) per
        batch entry after the last call; synchronises the stream."""
        rr = (ctypes.c_double * max(self._B, 1))()
        it = (ctypes.c_int * max(self._B, 1))()
        with torch.cuda.device(self._device):
            rc = self._lib.socu_mixed_stats(self._h, rr, it, ctypes.c_void_p(self._st().cuda_stream))
        _lib.check(rc, "cholesky_mixed_factor_and_solve stats")
        return np.array(rr[: self._B]), np.array(it[: self._B])

    def sweep_status(self):
        """(persistent, expired) of the FP32 correction solves, as for the
        regular launches.  Synchronises the stream."""
        p, e = ctypes.c_int(), ctypes.c_int()
        with torch.cuda.device(self._device):
            rc = self._lib.socu_mixed_sweep_status(self._h, ctypes.byref(p), ctypes.byref(e),
                                                   ctypes.c_void_p(self._st().cuda_stream))
        _lib.check(rc, "mixed sweep status")
        return bool(p.value), bool(e.value)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self._lib.socu_mixed_destroy(h)
            except Exception:
                pass
            self._h = ctypes.c_void_p()


def create_cholesky_mixed_factor_and_solve_launch(L, E, x, mask=None, precision="float32", tol=1e-10, max_iter=10,
                                                  device=None, stream=None, use_cuda_graph=False):
    """Mixed-precision factor_and_solve launch (see the module docstring).
    Arguments are validated like create_cholesky_factor_and_solve_launch; all
    arrays must be float64.  Returns a callable with a ``stats()`` method."""
    L, E, x, batch_dim, horizon, n, n_rhs = _batched(L, E, x)
    _check_mask(mask, batch_dim)
    if precision not in _PRECISIONS:
        raise ValueError(f"precision must be one of {sorted(_PRECISIONS)}")
    dev = _resolve_device(L, device, stream, use_cuda_graph)
    _check_tensors(torch.float64, dev, L=L, E=E, x=x, mask=mask)
    return _MixedLaunch(L, E, x, mask, batch_dim, horizon, n, n_rhs, dev, stream, precision, tol, max_iter,
                        use_cuda_graph)
