"""Block-level parity of the HIP factor and sweeps against the CPU oracle.

The reference's tests only compare solutions; here every output block is
checked: tril(L) of every node, every E index of every level segment
(including fill-in), the forward-sweep result y and the solution x, against
oracle.nd on the same seeded inputs.  float64 to 1e-11 relative (the factor
is unique, only summation order differs), float32 to 2e-4 relative.
"""
import numpy as np
import pytest

from oracle import nd

pytestmark = pytest.mark.gpu

CASES = [(1, 1), (1, 5), (3, 2), (4, 3), (8, 7), (5, 9), (16, 20), (33, 17), (40, 33), (64, 12), (96, 9), (128, 6)]


def _rel(a, b):
    d = np.linalg.norm((a - b).ravel())
    s = max(np.linalg.norm(b.ravel()), 1e-300)
    return d / s


@pytest.mark.parametrize("n,N", CASES)
@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_factor_blocks_match_oracle(n, N, dtype, device):
    import torch
    from socu_amd.block_tridiag_solver import (
        create_cholesky_backward_substitution_launch,
        create_cholesky_factor_launch,
        create_cholesky_forward_substitution_launch,
    )

    tol = 1e-11 if dtype == "float64" else 2e-4
    tdt = getattr(torch, dtype)
    B, m = 2, 3
    Ls, Es, bs = [], [], []
    for b in range(B):
        L0, E0, b0, _, _ = nd.readme_problem(n, N, m, alpha=10.0, seed=100 + b)
        Ls.append(L0)
        Es.append(E0)
        bs.append(b0)
    L0, E0, b0 = np.stack(Ls), np.stack(Es), np.stack(bs)
    Lr, Er, yr = L0.copy(), E0.copy(), b0.copy()
    nd.factor(Lr, Er)
    nd.forward(Lr, Er, yr)
    xr = yr.copy()
    nd.backward(Lr, Er, xr)

    L = torch.tensor(L0, dtype=tdt, device=device)
    E = torch.tensor(E0, dtype=tdt, device=device)
    x = torch.tensor(b0, dtype=tdt, device=device)
    create_cholesky_factor_launch(L, E, dtype=tdt)()
    Lg = np.tril(L.cpu().numpy().astype(np.float64))
    Eg = E.cpu().numpy().astype(np.float64)
    for b in range(B):
        assert _rel(Lg[b], np.tril(Lr[b])) < tol, f"tril(L) mismatch batch {b}"
        for k in range(Er.shape[1]):
            assert _rel(Eg[b, k], Er[b, k]) < tol or np.linalg.norm(Er[b, k]) == 0 and np.linalg.norm(Eg[b, k]) < tol, \
                f"E[{k}] mismatch batch {b}"
    create_cholesky_forward_substitution_launch(L, E, x, dtype=tdt)()
    yg = x.cpu().numpy().astype(np.float64)
    assert _rel(yg, yr) < tol * 10
    create_cholesky_backward_substitution_launch(L, E, x, dtype=tdt)()
    xg = x.cpu().numpy().astype(np.float64)
    assert _rel(xg, xr) < tol * 100


@pytest.mark.parametrize("n,N", [(8, 20), (32, 33)])
def test_factor_matches_dense_nd_cholesky(n, N, device):
    """tril(L) and every E block equal the blocks of the Cholesky factor of the
    ND-permuted dense matrix (the factor is unique)."""
    import torch
    from socu_amd.block_tridiag_solver import create_cholesky_factor_launch

    L0, E0, _, _, _ = nd.readme_problem(n, N, 1, alpha=10.0, seed=7)
    _, _, Lx, Ex = nd.dense_nd_factor(L0, E0)
    L = torch.tensor(L0, device=device)
    E = torch.tensor(E0, device=device)
    create_cholesky_factor_launch(L, E)()
    assert _rel(np.tril(L.cpu().numpy()), Lx) < 1e-11
    assert _rel(E.cpu().numpy(), Ex) < 1e-11


def test_factor_is_deterministic(device):
    """Owner-computes kernels have no atomics: two runs are bitwise equal."""
    import torch
    from socu_amd.block_tridiag_solver import create_cholesky_factor_and_solve_launch

    L0, E0, b0, _, _ = nd.readme_problem(64, 40, 4, alpha=10.0, seed=3)
    outs = []
    for _ in range(2):
        L = torch.tensor(L0, device=device)
        E = torch.tensor(E0, device=device)
        x = torch.tensor(b0, device=device)
        create_cholesky_factor_and_solve_launch(L, E, x)()
        outs.append((np.tril(L.cpu().numpy()), E.cpu().numpy(), x.cpu().numpy()))
    for a, b in zip(outs[0], outs[1]):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("n", [100, 128])
@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_wide_level_factor_matches_oracle(n, dtype, device):
    """N=130, B=2: the widest levels (130 and 64 nodes over the batch) take the
    chunked trsm and Schur variants for large blocks; every tril(L) and E block
    against the oracle (alpha=16: the generator stays well conditioned at
    n >= 100)."""
    import torch
    from socu_amd.block_tridiag_solver import create_cholesky_factor_launch

    tol = 1e-11 if dtype == "float64" else 2e-4
    tdt = getattr(torch, dtype)
    N, B = 130, 2
    Ls, Es = [], []
    for b in range(B):
        L0, E0, _, _, _ = nd.readme_problem(n, N, 1, alpha=16.0, seed=500 + b)
        Ls.append(L0)
        Es.append(E0)
    L0, E0 = np.stack(Ls), np.stack(Es)
    Lr, Er = L0.copy(), E0.copy()
    nd.factor(Lr, Er)
    L = torch.tensor(L0, dtype=tdt, device=device)
    E = torch.tensor(E0, dtype=tdt, device=device)
    create_cholesky_factor_launch(L, E, dtype=tdt)()
    Lg = np.tril(L.cpu().numpy().astype(np.float64))
    Eg = E.cpu().numpy().astype(np.float64)
    for b in range(B):
        for i in range(N):
            assert _rel(Lg[b, i], np.tril(Lr[b, i])) < tol, (b, i)
        for k in range(Er.shape[1]):
            if np.linalg.norm(Er[b, k]) > 0:
                assert _rel(Eg[b, k], Er[b, k]) < tol, (b, k)


@pytest.mark.parametrize("n,N,m", [(64, 20, 16), (64, 20, 17), (40, 13, 33), (128, 9, 33)])
@pytest.mark.parametrize("op", ["factor_and_solve", "factor+solve"])
def test_many_right_hand_sides_match_oracle(n, N, m, op, device):
    """Right-hand sides beyond one 16-wide chunk (16, 17, 33): the sweep
    kernels' chunk grid and a partial last chunk; E, tril(L) and x vs the
    oracle, and the residual."""
    import torch
    from socu_amd.block_tridiag_solver import (
        create_cholesky_factor_and_solve_launch,
        create_cholesky_factor_launch,
        create_cholesky_solve_launch,
    )

    L0, E0, b0, _, _ = nd.readme_problem(n, N, m, alpha=16.0, seed=21)
    Lr, Er, xr = L0.copy(), E0.copy(), b0.copy()
    nd.factor_and_solve(Lr, Er, xr)
    L = torch.tensor(L0, device=device)
    E = torch.tensor(E0, device=device)
    x = torch.tensor(b0, device=device)
    if op == "factor_and_solve":
        create_cholesky_factor_and_solve_launch(L, E, x, use_cuda_graph=True)()
    else:
        create_cholesky_factor_launch(L, E)()
        create_cholesky_solve_launch(L, E, x)()
    assert _rel(np.tril(L.cpu().numpy()), np.tril(Lr)) < 1e-11
    assert _rel(E.cpu().numpy(), Er) < 1e-11
    xg = x.cpu().numpy()
    assert _rel(xg, xr) < 1e-11
    assert nd.rel_residual(L0, E0, xg, b0) <= 1e-12
