"""The allocation-free solve entry point, socu_cholesky_solve_noalloc,
against socu_cholesky_solve with a caller-provided workspace.

Both run the same kernels on the same factor, so their solutions must agree
bit for bit.  Two calls in a row, the second with a smaller system, check
that the shared scratch array carries nothing from one call into the next;
a size beyond the scratch must be refused before any device work.  Every
call goes through the C ABI (ctypes), and the factor comes from the oracle.
"""
import ctypes

import numpy as np
import pytest

from oracle import nd

pytestmark = pytest.mark.gpu


def _problem(n, N, m, B, seed):
    Ls, Es, bs = [], [], []
    for k in range(B):
        L0, E0, b0, _, _ = nd.readme_problem(n, N, m, alpha=16.0, seed=seed + k)
        Ls.append(L0)
        Es.append(E0)
        bs.append(b0)
    L0, E0, b0 = np.stack(Ls), np.stack(Es), np.stack(bs)
    Lf, Ef = L0.copy(), E0.copy()
    nd.factor(Lf, Ef)
    xr = b0.copy()
    nd.forward(Lf, Ef, xr)
    nd.backward(Lf, Ef, xr)
    return Lf, Ef, b0, xr


def _solve_both(lib, Lf, Ef, b0, dtype, device):
    import torch

    from socu_amd import _lib

    B, N, n, m = b0.shape
    code = _lib.DTYPE_F64 if dtype == torch.float64 else _lib.DTYPE_F32
    L = torch.tensor(Lf, dtype=dtype, device=device)
    E = torch.tensor(Ef, dtype=dtype, device=device)
    x_ref = torch.tensor(b0, dtype=dtype, device=device)
    x_na = x_ref.clone()
    work = torch.empty(lib.socu_solve_workspace_size(B, N, n, code), dtype=torch.uint8, device=device)
    st = ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
    _lib.check(lib.socu_cholesky_solve(L.data_ptr(), E.data_ptr(), x_ref.data_ptr(), None, B, N, n, m, code,
                                       work.data_ptr(), st), "solve")
    _lib.check(lib.socu_cholesky_solve_noalloc(L.data_ptr(), E.data_ptr(), x_na.data_ptr(), None, B, N, n, m, code,
                                               st), "solve_noalloc")
    torch.cuda.synchronize(device)
    return x_ref.cpu().numpy(), x_na.cpu().numpy()


@pytest.mark.parametrize("dtype_name", ["float64", "float32"])
@pytest.mark.parametrize("first,second", [((128, 64, 1, 2), (32, 20, 3, 1)), ((64, 100, 16, 1), (10, 7, 1, 3))])
def test_noalloc_solve_matches_workspace_solve_twice(first, second, dtype_name, device):
    import torch

    from socu_amd import _lib

    dtype = getattr(torch, dtype_name)
    lib = _lib.load()
    for seed, (n, N, m, B) in ((11, first), (23, second)):
        print(f"noalloc case n={n} N={N} m={m} B={B} {dtype_name}", flush=True)
        Lf, Ef, b0, xr = _problem(n, N, m, B, seed)
        x_ref, x_na = _solve_both(lib, Lf, Ef, b0, dtype, device)
        np.testing.assert_array_equal(x_na, x_ref)
        tol = 1e-11 if dtype == torch.float64 else 2e-4
        assert np.linalg.norm((x_na - xr).ravel()) / np.linalg.norm(xr.ravel()) < tol


def test_noalloc_solve_with_mask(device):
    import torch

    from socu_amd import _lib

    lib = _lib.load()
    Lf, Ef, b0, xr = _problem(32, 40, 2, 3, seed=5)
    L = torch.tensor(Lf, device=device)
    E = torch.tensor(Ef, device=device)
    x = torch.tensor(b0, device=device)
    mask = torch.tensor([1, 0, 1], dtype=torch.bool, device=device)
    st = ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
    _lib.check(lib.socu_cholesky_solve_noalloc(L.data_ptr(), E.data_ptr(), x.data_ptr(), mask.data_ptr(), 3, 40, 32,
                                               2, _lib.DTYPE_F64, st), "solve_noalloc")
    xg = x.cpu().numpy()
    np.testing.assert_array_equal(xg[1], b0[1])
    for k in (0, 2):
        assert np.linalg.norm(xg[k] - xr[k]) / np.linalg.norm(xr[k]) < 1e-11


def test_noalloc_solve_refuses_sizes_beyond_its_scratch(device):
    import torch

    from socu_amd import _lib

    lib = _lib.load()
    # 2 * 512 FP64 blocks of 128 x 128 = 128 MiB of inverses > 64 MiB
    B, N, n, m = 2, 512, 128, 1
    assert lib.socu_solve_workspace_size(B, N, n, _lib.DTYPE_F64) > (64 << 20)
    L = torch.zeros(1, dtype=torch.float64, device=device)
    st = ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
    rc = lib.socu_cholesky_solve_noalloc(L.data_ptr(), L.data_ptr(), L.data_ptr(), None, B, N, n, m, _lib.DTYPE_F64,
                                         st)
    assert rc == _lib.EUNSUPPORTED
    assert "scratch" in _lib.last_error()
