"""Block sizes above 128 (the tiled kernels, n <= 256) against the oracle.

The reference takes any n (socu/blas/potrf.py:51-94 and the tails of
trsm.py, syrk.py, gemm.py); here n in (128, 256] runs the tiled path, whose
blocks stay in HBM and pass through LDS 32 x 32 tiles at a time.  Checked as
for the LDS-resident sizes: tril(L) and every E block after the factor, y
after the forward sweep, x after the backward sweep and after the solve,
FP64 to 1e-11 relative, with and without a graph, with a
mask, several right-hand-side chunks (m = 33) and a factor made by the oracle
(solves read only tril(L)).  FP32 is held to 1e-3: its error grows with the
conditioning of the larger blocks.
"""
import numpy as np
import pytest

from oracle import nd

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300)


def _batch(n, N, m, B, seed):
    Ls, Es, bs = [], [], []
    for k in range(B):
        L0, E0, b0, _, _ = nd.readme_problem(n, N, m, alpha=16.0, seed=seed + k)
        Ls.append(L0)
        Es.append(E0)
        bs.append(b0)
    return np.stack(Ls), np.stack(Es), np.stack(bs)


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("n,N,m,graph", [(129, 7, 1, False), (160, 20, 3, True), (200, 3, 2, False),
                                         (256, 9, 1, True), (250, 1, 1, False), (140, 2, 33, False)])
def test_large_factor_and_sweeps_vs_oracle(n, N, m, graph, dtype, device):
    import torch
    from socu_amd.block_tridiag_solver import (
        create_cholesky_backward_substitution_launch,
        create_cholesky_factor_and_solve_launch,
        create_cholesky_factor_launch,
        create_cholesky_forward_substitution_launch,
        create_cholesky_solve_launch,
    )

    print(f"large n={n} N={N} m={m} graph={graph} {dtype}", flush=True)
    tdt = getattr(torch, dtype)
    # FP32: kappa(Psi) * 2^-24 grows with n (4e-4 relative at n = 256); the
    # FP64 cases pin the algorithm at 1e-11
    tol = 1e-11 if dtype == "float64" else 1e-3
    B = 2
    L0, E0, b0 = _batch(n, N, m, B, seed=3)
    Lr, Er = L0.copy(), E0.copy()
    nd.factor(Lr, Er)
    yr = b0.copy()
    nd.forward(Lr, Er, yr)
    xr = yr.copy()
    nd.backward(Lr, Er, xr)

    L = torch.tensor(L0, dtype=tdt, device=device)
    E = torch.tensor(E0, dtype=tdt, device=device)
    create_cholesky_factor_launch(L, E, dtype=tdt, use_cuda_graph=graph)()
    Lg, Eg = L.double().cpu().numpy(), E.double().cpu().numpy()
    assert _rel(np.tril(Lg), np.tril(Lr)) < tol
    assert _rel(Eg, Er) < tol
    x = torch.tensor(b0, dtype=tdt, device=device)
    create_cholesky_forward_substitution_launch(L, E, x, dtype=tdt, use_cuda_graph=graph)()
    assert _rel(x.double().cpu().numpy(), yr) < tol
    create_cholesky_backward_substitution_launch(L, E, x, dtype=tdt, use_cuda_graph=graph)()
    assert _rel(x.double().cpu().numpy(), xr) < tol
    x = torch.tensor(b0, dtype=tdt, device=device)
    create_cholesky_solve_launch(L, E, x, dtype=tdt, use_cuda_graph=graph)()
    assert _rel(x.double().cpu().numpy(), xr) < tol
    # factor_and_solve on fresh inputs
    L2 = torch.tensor(L0, dtype=tdt, device=device)
    E2 = torch.tensor(E0, dtype=tdt, device=device)
    x2 = torch.tensor(b0, dtype=tdt, device=device)
    create_cholesky_factor_and_solve_launch(L2, E2, x2, dtype=tdt, use_cuda_graph=graph)()
    assert _rel(x2.double().cpu().numpy(), xr) < tol
    assert _rel(np.tril(L2.double().cpu().numpy()), np.tril(Lr)) < tol


def test_large_foreign_factor_and_mask(device):
    """Solves of an oracle factor (noise above the diagonal) with a mask that
    leaves one system untouched."""
    import torch
    from socu_amd.block_tridiag_solver import create_cholesky_solve_launch

    n, N, m, B = 176, 12, 2, 3
    L0, E0, b0 = _batch(n, N, m, B, seed=11)
    Lf, Ef = L0.copy(), E0.copy()
    nd.factor(Lf, Ef)
    xr = b0.copy()
    nd.solve(Lf, Ef, xr)
    noise = np.triu(np.random.RandomState(2).randn(*Lf.shape) * 1e3, 1)
    L = torch.tensor(np.tril(Lf) + noise, device=device)
    E = torch.tensor(Ef, device=device)
    x = torch.tensor(b0, device=device)
    mask = torch.tensor([True, False, True], device=device)
    for graph in (False, True):
        x.copy_(torch.tensor(b0, device=device))
        create_cholesky_solve_launch(L, E, x, mask=mask, use_cuda_graph=graph)()
        xg = x.cpu().numpy()
        np.testing.assert_array_equal(xg[1], b0[1])
        for k in (0, 2):
            assert _rel(xg[k], xr[k]) < 1e-11


def test_large_residual_and_limits(device):
    """n = 256, N = 64: rel-resid of the FP64 factor_and_solve at the FP64
    level (README generator, alpha = 32); n = 257 and the mixed / FP32-factor entry points above 128 are
    refused before any device work."""
    import torch
    from socu_amd import _lib
    from socu_amd.block_tridiag_solver import create_cholesky_factor_and_solve_launch
    from socu_amd.mixed import create_cholesky_mixed_factor_and_solve_launch

    n, N = 256, 64
    # alpha = 32: at n = 256 the generator's tril(randn) + alpha I factors are
    # numerically singular for alpha = 16 (the oracle's solve is NaN there);
    # at 32 the oracle reaches rel-resid 4.4e-15
    L0, E0, b0, _, _ = nd.readme_problem(n, N, 1, alpha=32.0, seed=1)
    L = torch.tensor(L0, device=device)
    E = torch.tensor(E0, device=device)
    x = torch.tensor(b0, device=device)
    create_cholesky_factor_and_solve_launch(L, E, x, use_cuda_graph=True)()
    assert nd.rel_residual(L0, E0, x.cpu().numpy(), b0) < 1e-13
    lib = _lib.load()
    assert lib.socu_max_block_size() == 256
    Lb = torch.zeros(2, 257, 257, dtype=torch.float64, device=device)
    Eb = torch.zeros(1, 257, 257, dtype=torch.float64, device=device)
    with pytest.raises(NotImplementedError):
        create_cholesky_factor_and_solve_launch(Lb, Eb, torch.zeros(2, 257, 1, dtype=torch.float64, device=device))
    with pytest.raises(NotImplementedError):
        create_cholesky_mixed_factor_and_solve_launch(L, E, x)
