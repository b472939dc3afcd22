"""The reference's own test grid, run against the HIP path.

Mirrors PREDICT-EPFL/socu tests/test_block_tridiag_solver.py: the parametric
solver grid (:132-140, body :54-130), the mask semantics (:142-193) and the
mask shape check (:195-201, here in test_api_cpu.py).  Data come from the same
builder (oracle.nd.reference_test_problem restates :17-52) and the reference
solution from scipy's dense Cholesky, as in :47-50; tolerances are the
reference's absolute 2-norms (:9-12).
"""
import functools

import numpy as np
import pytest
import scipy.linalg as la

from oracle import nd

pytestmark = pytest.mark.gpu

TOLERANCE = {"float32": 1e-4, "float64": 1e-10}


@functools.lru_cache(maxsize=64)
def prepare_test_data(n, N):
    A, L, E, b = nd.reference_test_problem(n, N)
    Lc = la.cholesky(A, lower=True)
    y = la.solve_triangular(Lc, b.reshape(-1), lower=True)
    x_ref = la.solve_triangular(Lc.T, y, lower=False)
    return A, L, E, b, x_ref


def _t(a, dtype, device):
    import torch

    return torch.tensor(a, dtype=getattr(torch, dtype), device=device)


def run_solver_test(n, N, use_cuda_graph, dtype, device):
    import torch
    from socu_amd.block_tridiag_solver import (
        create_cholesky_backward_substitution_launch,
        create_cholesky_factor_and_solve_launch,
        create_cholesky_factor_launch,
        create_cholesky_forward_substitution_launch,
        create_cholesky_solve_launch,
    )

    tol = TOLERANCE[dtype]
    tdt = getattr(torch, dtype)
    A, L_np, E_np, b_np, x_ref = prepare_test_data(n, N)
    L = _t(L_np, dtype, device)
    E = _t(E_np, dtype, device)
    x = _t(b_np, dtype, device)
    kw = dict(device=device, use_cuda_graph=use_cuda_graph, dtype=tdt)
    factor = create_cholesky_factor_launch(L, E, **kw)
    solve = create_cholesky_solve_launch(L, E, x, **kw)
    factor_and_solve = create_cholesky_factor_and_solve_launch(L, E, x, **kw)
    forward = create_cholesky_forward_substitution_launch(L, E, x, **kw)
    backward = create_cholesky_backward_substitution_launch(L, E, x, **kw)

    def reset(all_=True):
        if all_:
            L.copy_(_t(L_np, dtype, device))
            E.copy_(_t(E_np, dtype, device))
        x.copy_(_t(b_np, dtype, device))

    tag = f"n={n}, N={N}, dtype={dtype}, graph={use_cuda_graph}"
    reset()
    factor()
    solve()
    x_sep = x.cpu().numpy().astype(np.float64).reshape(-1)
    err = la.norm(x_ref - x_sep)
    res = la.norm(b_np.reshape(-1) - A @ x_sep)
    assert err < tol, f"solution error {err} for {tag}"
    assert res < tol, f"residual error {res} for {tag}"

    reset()
    factor_and_solve()
    x_comb = x.cpu().numpy().astype(np.float64).reshape(-1)
    err = la.norm(x_ref - x_comb)
    res = la.norm(b_np.reshape(-1) - A @ x_comb)
    assert err < tol, f"combined solution error {err} for {tag}"
    assert res < tol, f"combined residual error {res} for {tag}"
    assert la.norm(x_sep - x_comb) < tol, f"separate vs combined mismatch for {tag}"

    # L, E hold the factor from factor_and_solve
    reset(all_=False)
    forward()
    backward()
    x_split = x.cpu().numpy().astype(np.float64).reshape(-1)
    assert la.norm(x_sep - x_split) < tol, f"solve vs forward+backward mismatch for {tag}"


@pytest.mark.parametrize("N", [1, 2, 3, 4, 7, 8, 20, 100])
@pytest.mark.parametrize("use_cuda_graph", [False, True])
@pytest.mark.parametrize("n", [1, 10, 32, 36, 64, 68])
@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_cholesky_solver(n, N, use_cuda_graph, dtype, device):
    run_solver_test(n, N, use_cuda_graph, dtype, device)


@pytest.mark.parametrize("N", [1, 7, 20])
@pytest.mark.parametrize("use_cuda_graph", [False, True])
@pytest.mark.parametrize("n", [10, 36])
@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_cholesky_solver_mask(n, N, use_cuda_graph, dtype, device):
    """Active batch entries are solved; inactive entries stay bit-identical,
    also when the mask changes in place under a captured graph."""
    import torch
    from socu_amd.block_tridiag_solver import (
        create_cholesky_backward_substitution_launch,
        create_cholesky_factor_and_solve_launch,
        create_cholesky_factor_launch,
        create_cholesky_forward_substitution_launch,
        create_cholesky_solve_launch,
    )

    B = 4
    _, L_np, E_np, b_np, x_ref = prepare_test_data(n, N)
    L_np = np.tile(L_np, (B, 1, 1, 1))
    E_np = np.tile(E_np, (B, 1, 1, 1))
    b_np = np.tile(b_np.reshape(N, n, 1), (B, 1, 1, 1))
    tdt = getattr(torch, dtype)
    L = _t(L_np, dtype, device)
    E = _t(E_np, dtype, device)
    x = _t(b_np, dtype, device)
    mask = torch.zeros(B, dtype=torch.bool, device=device)
    kw = dict(device=device, use_cuda_graph=use_cuda_graph, dtype=tdt)
    factor = create_cholesky_factor_launch(L, E, mask, **kw)
    solve = create_cholesky_solve_launch(L, E, x, mask, **kw)
    factor_and_solve = create_cholesky_factor_and_solve_launch(L, E, x, mask, **kw)
    forward = create_cholesky_forward_substitution_launch(L, E, x, mask, **kw)
    backward = create_cholesky_backward_substitution_launch(L, E, x, mask, **kw)
    sequences = {
        "factor + solve": [factor, solve],
        "factor_and_solve": [factor_and_solve],
        "factor + forward + backward": [factor, forward, backward],
    }
    for mask_np in ([True, False, True, False], [False, True, False, True]):
        mask.copy_(torch.tensor(mask_np, device=device))
        for name, launches in sequences.items():
            for arr, arr_np in ((L, L_np), (E, E_np), (x, b_np)):
                arr.copy_(_t(arr_np, dtype, device))
            for launch in launches:
                launch()
            outs = ((L.cpu().numpy(), L_np), (E.cpu().numpy(), E_np), (x.cpu().numpy(), b_np))
            for k, active in enumerate(mask_np):
                label = f"{name}, mask={mask_np}, batch {k}"
                if active:
                    err = la.norm(x_ref - outs[2][0][k].astype(np.float64).reshape(-1))
                    assert err < TOLERANCE[dtype], f"{label}: error too large: {err}"
                else:
                    for out, ref in outs:
                        np.testing.assert_array_equal(out[k], ref[k].astype(out.dtype),
                                                      err_msg=f"{label}: inactive entry modified")
