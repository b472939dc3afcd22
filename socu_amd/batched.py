"""Batched launches over many independent systems of uneven size.

The reference batches only systems that share the horizon N and block size n
(``socu/block_tridiag_solver.py:749-768`` reshapes every operand to a common
``(B, N, n, n)``); MPC workloads hold thousands of systems of different N and
n.  ``create_batched_cholesky_launch`` takes a list of systems, each in the
reference's own layout (``L (N,n,n)``, ``E (calculate_off_diag_storage_len(N),
n, n)``, whose fill-in region the factor overwrites, ``x (N,n,m)``), and runs the chosen
operation on all of them in one graph-capturable launch set.  The per-system
result is the reference's factor / solve of that system
(``block_tridiag_solver.py:739-1727``), in place, in the same storage.
"""
from __future__ import annotations

import ctypes
from typing import Callable, Optional, Sequence

import torch

from . import _lib
from .block_tridiag_solver import _check_tensors, _dtype_name, _resolve_device, calculate_off_diag_storage_len

_OPS = {
    "factor": _lib.OP_FACTOR,
    "solve": _lib.OP_SOLVE,
    "factor_and_solve": _lib.OP_FACTOR_AND_SOLVE,
    "forward_substitution": _lib.OP_FORWARD,
    "backward_substitution": _lib.OP_BACKWARD,
}


def _system_shape(k, L, E, x, need_x):
    if L.ndim != 3 or L.shape[1] != L.shape[2]:
        raise ValueError(f"system {k}: L must have shape (N, n, n), got {tuple(L.shape)}")
    N, n = int(L.shape[0]), int(L.shape[1])
    want_e = (calculate_off_diag_storage_len(N), n, n)
    if tuple(E.shape) != want_e:
        raise ValueError(f"system {k}: E must have shape {want_e}, got {tuple(E.shape)}")
    m = 1
    if need_x:
        if x is None:
            raise ValueError(f"system {k}: this operation needs a right-hand side x")
        if x.ndim != 3 or x.shape[0] != N or x.shape[1] != n:
            raise ValueError(f"system {k}: x must have shape ({N}, {n}, m), got {tuple(x.shape)}")
        m = int(x.shape[2])
    return N, n, m


class BatchedLaunch:
    """Callable returned by :func:`create_batched_cholesky_launch`."""

    def __init__(self, op_code, systems, mask, code, device, stream, tune, use_graph, name):
        self._lib = _lib.load()
        self._keep = (systems, mask)  # the plan holds raw pointers into these
        self._device = device
        self._stream = stream
        self._name = name
        self._count = len(systems)
        table = (_lib.BatchSystem * max(1, len(systems)))()
        for k, (L, E, x, N, n, m) in enumerate(systems):
            table[k] = _lib.BatchSystem(L.data_ptr(), E.data_ptr(), None if x is None else x.data_ptr(), N, n, m, 0)
        self._plan = ctypes.c_void_p()
        with torch.cuda.device(device):
            rc = self._lib.socu_batched_create(ctypes.byref(self._plan), op_code, table, len(systems),
                                               None if mask is None else mask.data_ptr(), code, int(bool(tune)),
                                               int(bool(use_graph)))
        _lib.check(rc, name)

    def __call__(self) -> None:
        st = self._stream if self._stream is not None else torch.cuda.current_stream(self._device)
        with torch.cuda.device(self._device):
            rc = self._lib.socu_batched_execute(self._plan, ctypes.c_void_p(st.cuda_stream))
        _lib.check(rc, self._name)

    def config(self) -> dict:
        """Per size bucket (max(n, m) <= 8, 16, 32): the number of systems, the
        tile size and the wavefronts per workgroup they run with (as chosen
        by the autotuner when enabled)."""
        counts = (ctypes.c_int * 3)()
        tiles = (ctypes.c_int * 3)()
        waves = (ctypes.c_int * 3)()
        _lib.check(self._lib.socu_batched_config(self._plan, counts, tiles, waves), self._name)
        return {bound: {"systems": counts[b], "tile": tiles[b], "waves": waves[b]}
                for b, bound in enumerate((8, 16, 32))}

    def __del__(self):
        plan = getattr(self, "_plan", None)
        if plan is not None and plan.value:
            try:
                self._lib.socu_batched_destroy(plan)
            except Exception:
                pass
            self._plan = ctypes.c_void_p()


def create_batched_cholesky_launch(L: Sequence[torch.Tensor], E: Sequence[torch.Tensor],
                                   x: Optional[Sequence[torch.Tensor]] = None, op: str = "factor_and_solve",
                                   mask: Optional[torch.Tensor] = None, dtype=torch.float64, device=None, stream=None,
                                   use_cuda_graph: bool = False, tune: bool = True) -> Callable[[], None]:
    """Launch ``op`` ("factor", "solve", "factor_and_solve",
    "forward_substitution", "backward_substitution") on every system
    ``(L[k], E[k], x[k])``.  Systems may differ in N, n (<= 32) and m (<= 32).

    ``mask`` is a bool tensor with one entry per system; a False entry leaves
    that system untouched, as in the reference's mask (README.md:214-235).
    ``tune`` times every admissible tile size (8, 16, 32, at least the
    bucket's bound) and workgroup size (1, 2 or 4 wavefronts) for each size
    bucket on a device copy of the batch when the launch is created and keeps
    the fastest (cached per process, and in the file named by
    SOCU_TUNE_CACHE).
    """
    if op not in _OPS:
        raise ValueError(f"unknown op {op!r}; expected one of {sorted(_OPS)}")
    if _dtype_name(dtype) is None:
        raise ValueError(f"Unsupported dtype: {dtype}")
    op_code = _OPS[op]
    need_x = op != "factor"
    L, E = list(L), list(E)
    if len(E) != len(L):
        raise ValueError(f"got {len(L)} L blocks but {len(E)} E blocks")
    xs = list(x) if x is not None else [None] * len(L)
    if len(xs) != len(L):
        raise ValueError(f"got {len(L)} systems but {len(xs)} right-hand sides")
    if mask is not None and (mask.dtype != torch.bool or mask.ndim != 1 or mask.shape[0] != len(L)):
        raise ValueError(f"mask must be a bool tensor of shape ({len(L)},)")
    if not L:
        raise ValueError("the batch holds no systems")
    dev = _resolve_device(L[0], device, stream, use_cuda_graph)
    code = None
    systems = []
    for k, (Lk, Ek, xk) in enumerate(zip(L, E, xs)):
        code = _check_tensors(dtype, dev, L=Lk, E=Ek, x=xk if need_x else None)
        N, n, m = _system_shape(k, Lk, Ek, xk, need_x)
        if n > 32 or m > 32:
            raise NotImplementedError(f"system {k}: batched systems need n <= 32 and m <= 32 (got n={n}, m={m})")
        systems.append((Lk, Ek, xk if need_x else None, N, n, m))
    _check_tensors(dtype, dev, mask=mask)
    return BatchedLaunch(op_code, systems, mask, code, dev, stream, tune, use_cuda_graph,
                         f"create_batched_cholesky_launch({op})")
