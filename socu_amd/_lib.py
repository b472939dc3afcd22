"""ctypes binding of libsocu_hip.so (the C ABI declared in include/socu_hip.h).

torch is imported before the library is opened: torch ships a HIP runtime
whose SONAME (libamdhip64.so.7) is the one libsocu_hip.so needs, so loading in
this order makes both share one runtime, one device context and one set of
streams.  There is no fallback: if the library is missing the import fails.
"""
from __future__ import annotations

import ctypes
import os

import torch  # noqa: F401  (must be loaded first, see module docstring)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsocu_hip.so")

DTYPE_F32 = 0
DTYPE_F64 = 1

OK = 0
EINVAL = -1
EUNSUPPORTED = -2
EHIP = -3
EEXPIRED = -4

OP_FACTOR = 0
OP_SOLVE = 1
OP_FACTOR_AND_SOLVE = 2
OP_FORWARD = 3
OP_BACKWARD = 4

KERNEL_CLASSES = ("potrf", "trsm", "schur", "forward", "backward", "refine")

PREC_FP32 = 0
PREC_FP32_F32PIV = 1
PREC_BF16 = 2

_vp = ctypes.c_void_p
_i = ctypes.c_int


class BatchSystem(ctypes.Structure):
    """socu_batch_system: one independent system of a batched launch."""

    _fields_ = [("L", _vp), ("E", _vp), ("x", _vp), ("N", _i), ("n", _i), ("m", _i), ("reserved", _i)]


# name -> (restype, argtypes); every symbol declared in include/socu_hip.h
SIGNATURES = {
    "socu_calculate_recursive_iterations": (_i, [_i]),
    "socu_calculate_off_diag_storage_len": (ctypes.c_longlong, [_i]),
    "socu_max_block_size": (_i, []),
    "socu_set_sweep_spin_limit": (_i, [_i]),
    "socu_solve_workspace_size": (ctypes.c_size_t, [_i, _i, _i, _i]),
    "socu_cholesky_factor": (_i, [_vp, _vp, _vp, _i, _i, _i, _i, _vp]),
    "socu_cholesky_factor_f32": (_i, [_vp, _vp, _vp, _i, _i, _i, _i, _vp]),
    "socu_cholesky_forward_substitution": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _vp, _vp]),
    "socu_cholesky_backward_substitution": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _vp, _vp]),
    "socu_cholesky_solve": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _vp, _vp]),
    "socu_cholesky_solve_with_inverses": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _vp]),
    "socu_cholesky_solve_noalloc": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _vp]),
    "socu_cholesky_factor_and_solve": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _vp]),
    "socu_cholesky_factor_from": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _vp]),
    "socu_cholesky_factor_and_solve_from": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _vp]),
    "socu_plan_create": (_i, [ctypes.POINTER(_vp), _i, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _i]),
    "socu_plan_execute": (_i, [_vp, _vp]),
    "socu_plan_destroy": (_i, [_vp]),
    "socu_plan_sweep_status": (_i, [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i), _vp]),
    "socu_profile_op": (_i, [_i, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _vp, _vp,
                             ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)]),
    "socu_mixed_create": (_i, [ctypes.POINTER(_vp), _vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, ctypes.c_double, _i, _i]),
    "socu_mixed_execute": (_i, [_vp, _vp]),
    "socu_mixed_stats": (_i, [_vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int), _vp]),
    "socu_mixed_destroy": (_i, [_vp]),
    "socu_mixed_sweep_status": (_i, [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i), _vp]),
    "socu_mixed_profile": (_i, [_vp, _vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)]),
    "socu_batched_max_block_size": (_i, []),
    "socu_batched_create": (_i, [ctypes.POINTER(_vp), _i, ctypes.POINTER(BatchSystem), _i, _vp, _i, _i, _i]),
    "socu_batched_execute": (_i, [_vp, _vp]),
    "socu_batched_config": (_i, [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "socu_batched_destroy": (_i, [_vp]),
    "socu_adjoint_scatter": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _vp]),
    "socu_factor_adjoint_scatter": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _vp]),
    "socu_last_error": (ctypes.c_char_p, []),
}

_lib = None


class SocuError(RuntimeError):
    """A call into libsocu_hip.so returned an error code."""


class SweepExpiredError(SocuError):
    """A bounded dependency wait of a persistent substitution sweep expired
    in an earlier call of the launch; that call's solution holds NaN."""


def load():
    """Open libsocu_hip.so once and bind every C-ABI symbol."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"socu_amd: {LIB_PATH} is missing; build it with `python -c \"import __graft_entry__ as g; g.build()\"` "
            "(hipcc --offload-arch=gfx950).  There is no CPU fallback."
        )
    lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    msg = load().socu_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str):
    if rc != OK:
        msg = last_error()
        if rc == EINVAL:
            raise ValueError(f"{what}: {msg}")
        if rc == EUNSUPPORTED:
            raise NotImplementedError(f"{what}: {msg}")
        if rc == EEXPIRED:
            raise SweepExpiredError(f"{what}: {msg}")
        raise SocuError(f"{what}: {msg} (code {rc})")
