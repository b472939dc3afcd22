"""socu_amd: nested-dissection block-tridiagonal Cholesky for AMD Instinct
MI355X (gfx950), a drop-in for socu's launch API on PyTorch-ROCm tensors.

The public API lives in ``socu_amd.block_tridiag_solver`` under the same names
as socu's ``socu.block_tridiag_solver``.
"""
from .block_tridiag_solver import (  # noqa: F401
    calculate_off_diag_storage_len,
    calculate_recursive_iterations,
    create_cholesky_backward_substitution_launch,
    create_cholesky_factor_and_solve_launch,
    create_cholesky_factor_launch,
    create_cholesky_forward_substitution_launch,
    create_cholesky_solve_launch,
    optimal_problem_settings,
)
from .batched import create_batched_cholesky_launch  # noqa: F401
from .mixed import create_cholesky_mixed_factor_and_solve_launch  # noqa: F401

__version__ = "0.1.0"
