"""ctypes wrapper of the oracle's C restatement (oracle/nd_oracle.c).

TEST INFRASTRUCTURE ONLY (tests/ and bench.py's CPU-baseline leg).  Arrays are
float64 C-contiguous numpy arrays, updated in place, shaped like the product
API: L (B,N,n,n), E (B,n_E,n,n), x (B,N,n,m).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libnd_oracle.so")
_lib = None


def load(build_if_missing: bool = True):
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB) and build_if_missing:
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    lib = ctypes.CDLL(LIB)
    P = ctypes.c_void_p
    I = ctypes.c_int
    lib.ndo_factor.argtypes = [P, P, P, I, I, I]
    lib.ndo_forward.argtypes = [P, P, P, P, I, I, I, I]
    lib.ndo_backward.argtypes = [P, P, P, P, I, I, I, I]
    lib.ndo_factor_and_solve.argtypes = [P, P, P, P, I, I, I, I]
    lib.ndo_set_threads.argtypes = [I]
    lib.ndo_set_threads.restype = I
    lib.ndo_off_diag_storage_len.argtypes = [I]
    lib.ndo_off_diag_storage_len.restype = ctypes.c_longlong
    _lib = lib
    return lib


def _p(a):
    return None if a is None else a.ctypes.data


def _chk(*arrs):
    for a in arrs:
        if a is not None:
            assert a.dtype == np.float64 and a.flags.c_contiguous and a.ndim == 4


def _mask(mask):
    return None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)


def set_threads(t: int) -> int:
    return load().ndo_set_threads(t)


def factor(L, E, mask=None):
    _chk(L, E)
    mk = _mask(mask)
    B, N, n = L.shape[:3]
    load().ndo_factor(_p(L), _p(E), _p(mk), B, N, n)


def forward(L, E, x, mask=None):
    _chk(L, E, x)
    mk = _mask(mask)
    B, N, n, m = x.shape
    load().ndo_forward(_p(L), _p(E), _p(x), _p(mk), B, N, n, m)


def backward(L, E, x, mask=None):
    _chk(L, E, x)
    mk = _mask(mask)
    B, N, n, m = x.shape
    load().ndo_backward(_p(L), _p(E), _p(x), _p(mk), B, N, n, m)


def factor_and_solve(L, E, x, mask=None):
    _chk(L, E, x)
    mk = _mask(mask)
    B, N, n, m = x.shape
    load().ndo_factor_and_solve(_p(L), _p(E), _p(x), _p(mk), B, N, n, m)


# The product's C-ABI shape (include/socu_hip.h), restated on host arrays by
# the ndo_cholesky_* entry points: same argument lists as the socu_cholesky_*
# direct entry points, so one harness drives either library.
CAPI_ENTRY_POINTS = ("cholesky_factor", "cholesky_forward_substitution", "cholesky_backward_substitution",
                     "cholesky_solve", "cholesky_factor_and_solve", "cholesky_solve_workspace_size")


def capi():
    """The library with the ndo_cholesky_* argument types set from the
    product's own ctypes signatures (socu_amd._lib.SIGNATURES)."""
    from socu_amd import _lib as plib

    lib = load()
    for name in CAPI_ENTRY_POINTS:
        src = "socu_solve_workspace_size" if name == "cholesky_solve_workspace_size" else "socu_" + name
        res, args = plib.SIGNATURES[src]
        fn = getattr(lib, "ndo_" + name)
        fn.restype = res
        fn.argtypes = args
    return lib
