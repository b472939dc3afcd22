"""The C-ABI library loads and exports every function include/*.h declares.

No compute entry point is called (there is no GPU here); only the pure host
functions are exercised.
"""
import ctypes
import glob
import os
import re

import pytest

from socu_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(socu_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_the_entry_points():
    names = declared_functions()
    for must in ("socu_cholesky_factor", "socu_cholesky_solve", "socu_cholesky_factor_and_solve",
                 "socu_cholesky_forward_substitution", "socu_cholesky_backward_substitution",
                 "socu_calculate_off_diag_storage_len", "socu_plan_create", "socu_plan_execute",
                 "socu_plan_destroy", "socu_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in sorted(declared_functions()):
        assert hasattr(lib, name), f"{name} declared in include/ but not exported"


def test_ctypes_bindings_cover_the_header():
    assert set(_lib.SIGNATURES) == declared_functions()


def test_host_only_functions():
    lib = _lib.load()
    assert lib.socu_calculate_off_diag_storage_len(512) == 1013
    assert lib.socu_calculate_recursive_iterations(512) == 10
    assert lib.socu_max_block_size() >= 128
    assert lib.socu_calculate_off_diag_storage_len(0) < 0
    assert "horizon" in _lib.last_error()


def test_invalid_arguments_fail_before_device_work():
    lib = _lib.load()
    rc = lib.socu_cholesky_factor(None, None, None, 1, 5, 4, 7, None)
    assert rc == _lib.EINVAL and "dtype" in _lib.last_error()
    rc = lib.socu_cholesky_factor(None, None, None, 1, 5, 4, _lib.DTYPE_F64, None)
    assert rc == _lib.EINVAL and "NULL" in _lib.last_error()
    rc = lib.socu_cholesky_factor(ctypes.c_void_p(8), ctypes.c_void_p(8), None, 1, 5, 1000, _lib.DTYPE_F64, None)
    assert rc == _lib.EUNSUPPORTED
    plan = ctypes.c_void_p()
    rc = lib.socu_plan_create(ctypes.byref(plan), 9, None, None, None, None, 1, 5, 4, 1, _lib.DTYPE_F64, 0)
    assert rc == _lib.EINVAL
    # B = 0 is a valid no-op
    assert lib.socu_cholesky_factor(None, None, None, 0, 5, 4, _lib.DTYPE_F64, None) == 0


def test_substitutions_need_a_workspace():
    lib = _lib.load()
    assert lib.socu_solve_workspace_size(2, 512, 128, _lib.DTYPE_F64) == 2 * 512 * 128 * 128 * 8
    assert lib.socu_solve_workspace_size(1, 7, 10, _lib.DTYPE_F32) == 7 * 10 * 10 * 4
    assert lib.socu_solve_workspace_size(1, 0, 10, _lib.DTYPE_F32) == 0
    p = ctypes.c_void_p(8)
    for fn in (lib.socu_cholesky_solve, lib.socu_cholesky_forward_substitution,
               lib.socu_cholesky_backward_substitution):
        rc = fn(p, p, p, None, 1, 5, 4, 1, _lib.DTYPE_F64, None, None)
        assert rc == _lib.EINVAL and "work" in _lib.last_error()
        # B = 0: nothing to do, no workspace needed
        assert fn(None, None, None, None, 0, 5, 4, 1, _lib.DTYPE_F64, None, None) == 0


def test_spin_limit_setter_is_host_only():
    lib = _lib.load()
    assert lib.socu_set_sweep_spin_limit(-1) == 0
    assert lib.socu_set_sweep_spin_limit(0) == 0


def test_out_of_place_factor_rejects_overlapping_or_missing_sources():
    """The out-of-place factor entry points check their sources on the host
    (before any device work): missing sources and sources that overlap the
    outputs are EINVAL."""
    lib = _lib.load()
    f64 = _lib.DTYPE_F64
    B, N, n = 2, 5, 4
    nL = B * N * n * n * 8
    nE = B * lib.socu_calculate_off_diag_storage_len(N) * n * n * 8
    L, E = 1 << 32, (1 << 32) + (1 << 24)  # placeholder addresses: never dereferenced on these paths
    Ls, Es = (1 << 33), (1 << 33) + (1 << 24)
    assert nL < (1 << 24) and nE < (1 << 24)
    rc = lib.socu_cholesky_factor_from(None, Es, L, E, None, B, N, n, f64, None)
    assert rc == _lib.EINVAL and "L_src" in _lib.last_error()
    rc = lib.socu_cholesky_factor_from(Ls, None, L, E, None, B, N, n, f64, None)
    assert rc == _lib.EINVAL and "E_src" in _lib.last_error()
    rc = lib.socu_cholesky_factor_from(L + 8, Es, L, E, None, B, N, n, f64, None)
    assert rc == _lib.EINVAL and "overlap" in _lib.last_error()
    rc = lib.socu_cholesky_factor_and_solve_from(Ls, E + nE - 8, L, E, Ls + (1 << 20), None, B, N, n, 1, f64, None)
    assert rc == _lib.EINVAL and "overlap" in _lib.last_error()
    rc = lib.socu_cholesky_factor_and_solve_from(Ls, Es, L, E, None, None, B, N, n, 1, f64, None)
    assert rc == _lib.EINVAL and "x" in _lib.last_error()
