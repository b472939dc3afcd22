"""Drop-in mirror of socu's launch API (socu/block_tridiag_solver.py) on
PyTorch-ROCm tensors, backed by hand-written gfx950 HIP kernels.

Same names, the same argument meaning, the same validation order and the same
exception types as the reference:

* ``calculate_recursive_iterations``   block_tridiag_solver.py:12-15
* ``calculate_off_diag_storage_len``   :18-23
* ``optimal_problem_settings``         :26-60
* ``create_cholesky_factor_launch``                  :739-923
* ``create_cholesky_factor_and_solve_launch``        :926-1233
* ``create_cholesky_forward_substitution_launch``    :1236-1374
* ``create_cholesky_backward_substitution_launch``   :1377-1517
* ``create_cholesky_solve_launch``                   :1520-1727

Arrays are ``torch.Tensor`` on a ROCm GPU (device type ``"cuda"``), contiguous,
``torch.float64`` or ``torch.float32``: L (N,n,n) or (B,N,n,n), E (n_E,n,n) or
(B,n_E,n,n), x (N,n,m) or (B,N,n,m), mask (B,) ``torch.bool``.  They are bound
at creation and updated in place by every call of the returned callable; the
mask is re-read by the kernels on every call, also from a captured graph.

``block_dim`` and ``block_size`` are accepted for signature compatibility; the
HIP kernels choose their own workgroup and tile shapes.  ``use_cuda_graph``
captures the work into a hipGraph once and replays it.  A callable enqueues on
``stream`` if given, else on the current torch stream of the device at call
time; it never synchronises the host.  There is no CPU path: CPU tensors raise
``ValueError``.
"""
from __future__ import annotations

import ctypes
from typing import Callable, Optional

import torch

from . import _lib

__all__ = [
    "calculate_recursive_iterations",
    "calculate_off_diag_storage_len",
    "optimal_problem_settings",
    "create_cholesky_factor_launch",
    "create_cholesky_factor_and_solve_launch",
    "create_cholesky_forward_substitution_launch",
    "create_cholesky_backward_substitution_launch",
    "create_cholesky_solve_launch",
]


def calculate_recursive_iterations(horizon: int) -> int:
    """floor(log2(horizon)) + 1 (block_tridiag_solver.py:12-15)."""
    assert horizon > 0
    return int(horizon).bit_length()


def calculate_off_diag_storage_len(horizon: int) -> int:
    """Number of E blocks of the ND storage: sum_l (horizon // 2^l - 1)
    (block_tridiag_solver.py:18-23)."""
    iterations = calculate_recursive_iterations(horizon)
    return sum(horizon // (2 ** i) - 1 for i in range(iterations))


def _dtype_name(dtype) -> Optional[str]:
    if dtype in (torch.float64,):
        return "float64"
    if dtype in (torch.float32,):
        return "float32"
    name = getattr(dtype, "__name__", None) or getattr(dtype, "name", None) or (dtype if isinstance(dtype, str) else None)
    if name in ("float64", "double", "f64"):
        return "float64"
    if name in ("float32", "float", "f32"):
        return "float32"
    return None


def optimal_problem_settings(n: int, dtype):
    """The reference's heuristic table (block_tridiag_solver.py:26-60), kept for
    callers that read it (e.g. the padding multiples).  Raises ValueError for a
    dtype other than float32/float64."""
    name = _dtype_name(dtype)
    if name == "float64":
        factor_fuse_threshold, solve_fuse_threshold = 16, 32
        pad_multiple_fuse, pad_multiple_blocked = 2, 8
    elif name == "float32":
        factor_fuse_threshold, solve_fuse_threshold = 32, 32
        pad_multiple_fuse, pad_multiple_blocked = 4, 16
    else:
        raise ValueError("dtype must be float64 or float32")
    settings = {
        "block_dim": {"factor": 64, "solve": 64},
        "block_size": {
            "factor": 32 if n % 32 == 0 else 16,
            "solve": 32 if n % 32 == 0 else 16,
        },
        "pad_multiple": {
            "factor": pad_multiple_fuse if n < factor_fuse_threshold else pad_multiple_blocked,
            "solve": pad_multiple_fuse if n < solve_fuse_threshold else pad_multiple_blocked,
        },
    }
    if n < factor_fuse_threshold:
        settings["block_size"]["factor"] = factor_fuse_threshold
    if n < solve_fuse_threshold:
        settings["block_size"]["solve"] = solve_fuse_threshold
    return settings


# ---------------------------------------------------------------------------
# validation shared by the five factories
# ---------------------------------------------------------------------------

def _batched(L, E, x):
    """Shape checks of the reference in its order (:749-768, :937-960)."""
    if x is None:
        assert E.ndim == L.ndim, "L and E must have the same number of dimensions"
    else:
        assert E.ndim == L.ndim and x.ndim == L.ndim, "L, E, and x must have the same number of dimensions"
    if L.ndim == 4:
        batch_dim = L.shape[0]
    elif L.ndim == 3:
        batch_dim = 1
        L = L.reshape((1, *L.shape))
        E = E.reshape((1, *E.shape))
        if x is not None:
            x = x.reshape((1, *x.shape))
    else:
        raise ValueError("L must have 4 or 3 dimensions")
    if x is None:
        assert E.shape[0] == batch_dim, "L and E must have same batch dimension"
    else:
        assert E.shape[0] == batch_dim and x.shape[0] == batch_dim, "L, E, and x must have same batch dimension"
    horizon = L.shape[1]
    n = L.shape[2]
    assert L.shape[3] == n, "The last two dimensions of L must be the same"
    assert E.shape[2] == n and E.shape[3] == n, "L and E must have same block size"
    n_E = calculate_off_diag_storage_len(horizon)
    assert E.shape[1] == n_E, "E has incorrect size, use calculate_off_diag_storage_len to calculate the correct number of blocks"
    n_rhs = 1
    if x is not None:
        n_rhs = x.shape[3]
        assert x.shape[1] == horizon and x.shape[2] == n, "x has incorrect dimensions"
    return L, E, x, batch_dim, horizon, n, n_rhs


def _check_mask(mask, batch_dim):
    if mask is not None:
        assert mask.ndim == 1 and mask.shape[0] == batch_dim, "mask must have shape (batch_dim,)"
        assert mask.dtype == torch.bool, "mask must be a torch.bool tensor"


def _resolve_device(L, device, stream, use_cuda_graph):
    """Mirror of socu.utils.resolve_launch_device (utils.py:15-22): the stream's
    device wins, then the device argument, then the tensors' device."""
    if stream is not None:
        dev = stream.device
    elif device is not None:
        dev = torch.device(device)
    else:
        dev = L.device
    if dev.type != "cuda":
        if use_cuda_graph:
            raise ValueError("use_cuda_graph=True requires a GPU device")
        raise ValueError(
            f"socu_amd runs on an AMD Instinct GPU (torch device type 'cuda' under ROCm); got device '{dev}'. "
            "There is no CPU implementation."
        )
    if dev.index is None:
        dev = torch.device("cuda", torch.cuda.current_device())
    return dev


def _check_tensors(dtype, dev, **arrays):
    want = torch.float64 if _dtype_name(dtype) == "float64" else torch.float32
    for name, t in arrays.items():
        if t is None:
            continue
        if not isinstance(t, torch.Tensor):
            raise TypeError(f"{name} must be a torch.Tensor, got {type(t).__name__}")
        if name != "mask" and t.dtype != want:
            raise TypeError(f"{name} has dtype {t.dtype} but the launch was created for {want}")
        if t.device != dev:
            raise ValueError(f"{name} is on {t.device}, expected {dev}")
        if not t.is_contiguous():
            raise ValueError(f"{name} must be contiguous (the launch updates it in place)")
    return _lib.DTYPE_F64 if want == torch.float64 else _lib.DTYPE_F32


def _ptr(t) -> Optional[int]:
    return None if t is None else t.data_ptr()


class _Launch:
    """The callable a launch factory returns: a native plan bound to the arrays
    (and, with ``use_cuda_graph``, its instantiated hipGraph)."""

    def __init__(self, op, L, E, x, mask, B, N, n, m, code, device, stream, use_graph, name):
        self._lib = _lib.load()
        self._keep = (L, E, x, mask)  # the plan holds raw pointers into these
        self._device = device
        self._stream = stream
        self._name = name
        self._plan = ctypes.c_void_p()
        with torch.cuda.device(device):
            rc = self._lib.socu_plan_create(
                ctypes.byref(self._plan), op, _ptr(L), _ptr(E), _ptr(x), _ptr(mask), B, N, n, m, code, int(bool(use_graph))
            )
        _lib.check(rc, name)

    def __call__(self) -> None:
        st = self._stream if self._stream is not None else torch.cuda.current_stream(self._device)
        with torch.cuda.device(self._device):
            rc = self._lib.socu_plan_execute(self._plan, ctypes.c_void_p(st.cuda_stream))
        _lib.check(rc, self._name)

    def sweep_status(self):
        """(persistent, expired): whether the substitution sweeps run as one
        persistent launch (one right-hand side), and whether a bounded
        dependency wait expired in the last call.  Synchronises the stream."""
        st = self._stream if self._stream is not None else torch.cuda.current_stream(self._device)
        p, e = ctypes.c_int(), ctypes.c_int()
        with torch.cuda.device(self._device):
            rc = self._lib.socu_plan_sweep_status(self._plan, ctypes.byref(p), ctypes.byref(e),
                                                  ctypes.c_void_p(st.cuda_stream))
        _lib.check(rc, self._name)
        return bool(p.value), bool(e.value)

    def __del__(self):
        plan = getattr(self, "_plan", None)
        if plan is not None and plan.value:
            try:
                self._lib.socu_plan_destroy(plan)
            except Exception:
                pass
            self._plan = ctypes.c_void_p()


def _make(op, name, L, E, x, mask, dtype, device, stream, use_cuda_graph, block_dim, block_size) -> Callable[[], None]:
    L, E, x, batch_dim, horizon, n, n_rhs = _batched(L, E, x)
    _check_mask(mask, batch_dim)
    optimal_problem_settings(n, dtype)  # ValueError for an unsupported dtype (:38-39)
    if block_size is not None and int(block_size) <= 0:
        raise ValueError("block_size must be positive")
    if block_dim is not None and int(block_dim) <= 0:
        raise ValueError("block_dim must be positive")
    dev = _resolve_device(L, device, stream, use_cuda_graph)
    code = _check_tensors(dtype, dev, L=L, E=E, x=x, mask=mask)
    return _Launch(op, L, E, x, mask, batch_dim, horizon, n, n_rhs, code, dev, stream, use_cuda_graph, name)


def create_cholesky_factor_launch(L, E, mask=None, block_dim=None, block_size=None, dtype=torch.float64, device=None,
                                  stream=None, use_cuda_graph=False):
    """In-place ND Cholesky factor of the block-tridiagonal matrix (L, E)
    (block_tridiag_solver.py:739-923).  The fill-in region E[..., N-1:, :, :]
    receives the fill-in; the reference needs it zeroed before each call, here
    each fill-in block is written once and its contents on entry are ignored."""
    return _make(_lib.OP_FACTOR, "cholesky_factor", L, E, None, mask, dtype, device, stream, use_cuda_graph,
                 block_dim, block_size)


def create_cholesky_factor_and_solve_launch(L, E, x, mask=None, block_dim=None, block_size=None, dtype=torch.float64,
                                            device=None, stream=None, use_cuda_graph=False):
    """Factor then solve in place; x holds b on entry and the solution on exit
    (block_tridiag_solver.py:926-1233)."""
    return _make(_lib.OP_FACTOR_AND_SOLVE, "cholesky_factor_and_solve", L, E, x, mask, dtype, device, stream,
                 use_cuda_graph, block_dim, block_size)


def create_cholesky_forward_substitution_launch(L, E, x, mask=None, block_dim=None, block_size=None,
                                                dtype=torch.float64, device=None, stream=None, use_cuda_graph=False):
    """x <- Lp^{-1} P x with the factor held in (L, E) (block_tridiag_solver.py:1236-1374)."""
    return _make(_lib.OP_FORWARD, "cholesky_forward_substitution", L, E, x, mask, dtype, device, stream,
                 use_cuda_graph, block_dim, block_size)


def create_cholesky_backward_substitution_launch(L, E, x, mask=None, block_dim=None, block_size=None,
                                                 dtype=torch.float64, device=None, stream=None, use_cuda_graph=False):
    """x <- P^T Lp^{-T} x with the factor held in (L, E) (block_tridiag_solver.py:1377-1517)."""
    return _make(_lib.OP_BACKWARD, "cholesky_backward_substitution", L, E, x, mask, dtype, device, stream,
                 use_cuda_graph, block_dim, block_size)


def create_cholesky_solve_launch(L, E, x, mask=None, block_dim=None, block_size=None, dtype=torch.float64, device=None,
                                 stream=None, use_cuda_graph=False):
    """Forward then backward substitution in place (block_tridiag_solver.py:1520-1727)."""
    return _make(_lib.OP_SOLVE, "cholesky_solve", L, E, x, mask, dtype, device, stream, use_cuda_graph, block_dim,
                 block_size)
