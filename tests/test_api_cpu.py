"""Host-side behaviour of the launch API, runnable without a GPU.

Mirrors the reference's argument validation (block_tridiag_solver.py:749-774,
:937-966, :26-60; utils.py:15-22) and its mask shape test
(tests/test_block_tridiag_solver.py:195-201).  Every check here raises before
any device work is enqueued.
"""
import pytest
import torch

from socu_amd import block_tridiag_solver as bts
from socu_amd.block_tridiag_solver import (
    calculate_off_diag_storage_len,
    create_cholesky_backward_substitution_launch,
    create_cholesky_factor_and_solve_launch,
    create_cholesky_factor_launch,
    create_cholesky_forward_substitution_launch,
    create_cholesky_solve_launch,
    optimal_problem_settings,
)

XFACT = [
    create_cholesky_solve_launch,
    create_cholesky_factor_and_solve_launch,
    create_cholesky_forward_substitution_launch,
    create_cholesky_backward_substitution_launch,
]


def arrays(B=2, N=5, n=4, m=1, dtype=torch.float64, batched=True):
    nE = calculate_off_diag_storage_len(N)
    if batched:
        return (torch.zeros(B, N, n, n, dtype=dtype), torch.zeros(B, nE, n, n, dtype=dtype),
                torch.zeros(B, N, n, m, dtype=dtype))
    return torch.zeros(N, n, n, dtype=dtype), torch.zeros(nE, n, n, dtype=dtype), torch.zeros(N, n, m, dtype=dtype)


def test_mask_shape_check():
    """tests/test_block_tridiag_solver.py:195-201."""
    L, E, _ = arrays(B=2, N=5, n=4)
    bad_mask = torch.ones(3, dtype=torch.bool)
    with pytest.raises(AssertionError):
        create_cholesky_factor_launch(L, E, mask=bad_mask)


@pytest.mark.parametrize("factory", XFACT)
def test_mask_shape_check_x_factories(factory):
    L, E, x = arrays(B=2)
    with pytest.raises(AssertionError):
        factory(L, E, x, mask=torch.ones(3, dtype=torch.bool))
    with pytest.raises(AssertionError):
        factory(L, E, x, mask=torch.ones(2, dtype=torch.uint8))


def test_wrong_e_length():
    L, E, _ = arrays(B=2, N=5)
    with pytest.raises(AssertionError, match="calculate_off_diag_storage_len"):
        create_cholesky_factor_launch(L, E[:, :-1])


def test_ndim_checks():
    L, E, x = arrays()
    with pytest.raises(AssertionError):
        create_cholesky_factor_launch(L, E[0])
    with pytest.raises(ValueError):
        create_cholesky_factor_launch(L[0, 0], E[0, 0])
    with pytest.raises(AssertionError):
        create_cholesky_solve_launch(L, E, x[0])


def test_shape_checks():
    L, E, x = arrays(B=2, N=5, n=4, m=3)
    with pytest.raises(AssertionError):
        create_cholesky_factor_launch(L[:, :, :, :3], E)  # non-square blocks
    with pytest.raises(AssertionError):
        create_cholesky_factor_launch(L, E[:, :, :3, :3])  # block size mismatch
    with pytest.raises(AssertionError):
        create_cholesky_factor_launch(L[:1], E)  # batch mismatch
    with pytest.raises(AssertionError):
        create_cholesky_solve_launch(L, E, x[:, :4])  # horizon mismatch
    with pytest.raises(AssertionError):
        create_cholesky_solve_launch(L, E, x[:, :, :3])  # n mismatch
    with pytest.raises(AssertionError):
        create_cholesky_factor_launch(torch.zeros(1, 0, 4, 4), torch.zeros(1, 0, 4, 4))  # N = 0


def test_bad_dtype_raises_value_error():
    L, E, _ = arrays()
    with pytest.raises(ValueError):
        create_cholesky_factor_launch(L, E, dtype=torch.float16)
    with pytest.raises(ValueError):
        optimal_problem_settings(32, torch.int32)


def test_cpu_tensors_are_rejected_without_fallback():
    L, E, x = arrays()
    with pytest.raises(ValueError, match="no CPU implementation"):
        create_cholesky_factor_launch(L, E)
    with pytest.raises(ValueError, match="use_cuda_graph"):
        create_cholesky_solve_launch(L, E, x, use_cuda_graph=True)


def test_three_dim_inputs_validate_like_batch_of_one():
    L, E, x = arrays(batched=False)
    with pytest.raises(ValueError, match="no CPU implementation"):
        create_cholesky_factor_and_solve_launch(L, E, x)


@pytest.mark.parametrize("n", [1, 8, 10, 16, 24, 31, 32, 36, 64, 68, 128])
def test_optimal_problem_settings_table(n):
    """Values of block_tridiag_solver.py:26-60."""
    f64 = optimal_problem_settings(n, torch.float64)
    f32 = optimal_problem_settings(n, "float32")
    bs = 32 if n % 32 == 0 else 16
    assert f64["block_dim"] == {"factor": 64, "solve": 64}
    assert f64["block_size"]["factor"] == (16 if n < 16 else bs)
    assert f64["block_size"]["solve"] == (32 if n < 32 else bs)
    assert f64["pad_multiple"] == {"factor": 2 if n < 16 else 8, "solve": 2 if n < 32 else 8}
    assert f32["block_size"]["factor"] == (32 if n < 32 else bs)
    assert f32["pad_multiple"] == {"factor": 4 if n < 32 else 16, "solve": 4 if n < 32 else 16}


def test_public_names_match_reference_api():
    for name in ("calculate_recursive_iterations", "calculate_off_diag_storage_len", "optimal_problem_settings",
                 "create_cholesky_factor_launch", "create_cholesky_solve_launch",
                 "create_cholesky_factor_and_solve_launch", "create_cholesky_forward_substitution_launch",
                 "create_cholesky_backward_substitution_launch"):
        assert callable(getattr(bts, name))
