"""factor_and_solve against the oracle, with the n = 32 fused kernels.

At n = 32 one kernel per level factors each node's diagonal block, applies
the coupling trsm and forms the fill-in block (the reference's potrf, trsm
and gemm passes, socu/blas/potrf.py:97-119, socu/blas/trsm.py:94-120 and
:202-228, block_tridiag_solver.py:797-813), and a factor_and_solve takes the
forward sweep inside the factor's kernels (block_tridiag_solver.py:926-1233
interleaves them likewise); other n run the level kernels with the forward
sweep beside the factor.  Checked for several n: tril(L) of every block,
every E index (fill-in blocks included, and the same blocks zero as the
oracle's) and x, FP64 to 1e-11 and FP32 to 5e-4 relative (the oracle in
FP64); masked systems untouched.  Right-hand-side widths cross the kernels'
column chunks (8 in the Schur kernel, 32 in the node kernel, 16 in the sweep
kernels); both the stateless C entry point and the launch object (captured
in a graph) are used.
"""
import numpy as np
import pytest

from oracle import nd

pytestmark = pytest.mark.gpu

TOL = {"float64": 1e-11, "float32": 5e-4}


def _rel(a, b):
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300)


def _stack(n, N, m, B, seed):
    out = [nd.readme_problem(n, N, m, alpha=16.0, seed=seed + k)[:3] for k in range(B)]
    return tuple(np.stack([o[i] for o in out]) for i in range(3))


def _check(Lg, Eg, xg, L0, E0, b0, mask, dtype):
    Lr, Er, xr = L0.copy(), E0.copy(), b0.copy()
    nd.factor_and_solve(Lr, Er, xr)
    tol = TOL[dtype]
    for k, on in enumerate(mask):
        if not on:
            np.testing.assert_array_equal(Lg[k], L0[k].astype(dtype))
            np.testing.assert_array_equal(Eg[k], E0[k].astype(dtype))
            np.testing.assert_array_equal(xg[k], b0[k].astype(dtype))
            continue
        assert _rel(np.tril(Lg[k]).astype(np.float64), np.tril(Lr[k])) < tol
        zr = np.all(Er[k] == 0, axis=(1, 2))
        np.testing.assert_array_equal(np.all(Eg[k] == 0, axis=(1, 2)), zr)
        for q in np.where(~zr)[0]:
            assert _rel(Eg[k][q].astype(np.float64), Er[k][q]) < tol, (k, q)
        assert _rel(xg[k].astype(np.float64), xr[k]) < tol


CASES = [(32, 1, 3), (32, 2, 1), (32, 3, 9), (32, 33, 40), (32, 64, 1), (32, 100, 9), (32, 17, 33),
         (8, 20, 3), (20, 37, 1), (40, 33, 2), (64, 64, 1), (64, 21, 17), (96, 17, 3), (100, 12, 9), (128, 40, 1),
         (128, 19, 10), (128, 2, 1)]


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("n,N,m", CASES)
def test_factor_and_solve_entry_point(n, N, m, dtype, device):
    import torch

    from socu_amd import _lib

    lib = _lib.load()
    code = _lib.DTYPE_F64 if dtype == "float64" else _lib.DTYPE_F32
    tdt = getattr(torch, dtype)
    B = 3
    L0, E0, b0 = _stack(n, N, m, B, seed=11 * N + m + n)
    L = torch.tensor(L0, dtype=tdt, device=device)
    E = torch.tensor(E0, dtype=tdt, device=device)
    x = torch.tensor(b0, dtype=tdt, device=device)
    mask = torch.tensor([1, 0, 1], dtype=torch.uint8, device=device)
    st = torch.cuda.current_stream(device).cuda_stream
    _lib.check(lib.socu_cholesky_factor_and_solve(L.data_ptr(), E.data_ptr(), x.data_ptr(), mask.data_ptr(), B, N, n,
                                                  m, code, st), "factor_and_solve")
    torch.cuda.synchronize()
    _check(L.cpu().numpy(), E.cpu().numpy(), x.cpu().numpy(), L0, E0, b0, [1, 0, 1], dtype)


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("n,N,m", [(32, 40, 4), (32, 9, 12), (128, 64, 1), (64, 30, 3)])
def test_factor_and_solve_launch_graph(n, N, m, dtype, device):
    """The launch object replayed three times from restored inputs."""
    import torch

    from socu_amd.block_tridiag_solver import create_cholesky_factor_and_solve_launch

    tdt = getattr(torch, dtype)
    B = 2
    L0, E0, b0 = _stack(n, N, m, B, seed=5 * N + m)
    L = torch.tensor(L0, dtype=tdt, device=device)
    E = torch.tensor(E0, dtype=tdt, device=device)
    x = torch.tensor(b0, dtype=tdt, device=device)
    launch = create_cholesky_factor_and_solve_launch(L, E, x, dtype=tdt, use_cuda_graph=True)
    for _ in range(3):
        L.copy_(torch.tensor(L0, dtype=tdt, device=device))
        E.copy_(torch.tensor(E0, dtype=tdt, device=device))
        x.copy_(torch.tensor(b0, dtype=tdt, device=device))
        launch()
        torch.cuda.synchronize()
        _check(L.cpu().numpy(), E.cpu().numpy(), x.cpu().numpy(), L0, E0, b0, [1] * B, dtype)
