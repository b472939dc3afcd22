"""Drop-in contract of the substitution launches and the failure path of the
persistent sweeps.

* The reference's solve, forward and backward substitution read only tril(L)
  and E (socu/blas/potrf.py:101-103 "Only the lower triangular part is
  used"): a factor made anywhere -- here by the oracle -- with arbitrary
  values in the strict upper triangle of each L[i] must give the oracle's y
  and x (FP64 to 1e-11 relative), and the launches must leave L and E
  bit-identical.
* A bounded dependency wait of the persistent sweeps that expires must show:
  the solution holds NaN and the next call of the launch raises.  The waits
  are forced to expire from the host (socu_set_sweep_spin_limit(-1): every
  wait that is not already satisfied gives up at once), so every kernel
  still exits normally.
"""
import ctypes

import numpy as np
import pytest

from oracle import nd

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300)


def _stack(n, N, m, B, seed, alpha=16.0):
    Ls, Es, bs = [], [], []
    for k in range(B):
        L0, E0, b0, _, _ = nd.readme_problem(n, N, m, alpha=alpha, seed=seed + k)
        Ls.append(L0)
        Es.append(E0)
        bs.append(b0)
    return np.stack(Ls), np.stack(Es), np.stack(bs)


@pytest.mark.parametrize("n,N,m", [(8, 20, 1), (32, 33, 1), (36, 17, 3), (64, 40, 1), (100, 9, 2), (128, 24, 1),
                                   (128, 16, 17)])
@pytest.mark.parametrize("graph", [False, True])
def test_substitutions_of_a_foreign_factor(n, N, m, graph, device):
    import torch
    from socu_amd.block_tridiag_solver import (
        create_cholesky_backward_substitution_launch,
        create_cholesky_forward_substitution_launch,
        create_cholesky_solve_launch,
    )

    B = 2
    L0, E0, b0 = _stack(n, N, m, B, seed=40)
    Lf, Ef = L0.copy(), E0.copy()
    nd.factor(Lf, Ef)
    yr = b0.copy()
    nd.forward(Lf, Ef, yr)
    xr = yr.copy()
    nd.backward(Lf, Ef, xr)
    # the oracle's factor in the lower triangle, noise above it
    rs = np.random.RandomState(5)
    noise = np.triu(rs.randn(*Lf.shape) * 1e3, 1)
    Lup = np.tril(Lf) + noise
    L = torch.tensor(Lup, device=device)
    E = torch.tensor(Ef, device=device)

    x = torch.tensor(b0, device=device)
    create_cholesky_forward_substitution_launch(L, E, x, use_cuda_graph=graph)()
    assert _rel(x.cpu().numpy(), yr) < 1e-11
    create_cholesky_backward_substitution_launch(L, E, x, use_cuda_graph=graph)()
    assert _rel(x.cpu().numpy(), xr) < 1e-11
    x.copy_(torch.tensor(b0, device=device))
    solve = create_cholesky_solve_launch(L, E, x, use_cuda_graph=graph)
    solve()
    assert _rel(x.cpu().numpy(), xr) < 1e-11
    # a second call on a fresh right-hand side (graph replay)
    x.copy_(torch.tensor(b0, device=device))
    solve()
    assert _rel(x.cpu().numpy(), xr) < 1e-11
    np.testing.assert_array_equal(L.cpu().numpy(), Lup)
    np.testing.assert_array_equal(E.cpu().numpy(), Ef)


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_solve_after_factor_of_this_library_then_tril(dtype, device):
    """factor with this library, clear the strict upper triangle (L.tril_()),
    then solve: the result does not depend on what the factor left above the
    diagonal."""
    import torch
    from socu_amd.block_tridiag_solver import create_cholesky_factor_launch, create_cholesky_solve_launch

    tdt = getattr(torch, dtype)
    n, N, m = 64, 33, 1
    L0, E0, b0 = _stack(n, N, m, 1, seed=3)
    L = torch.tensor(L0, dtype=tdt, device=device)
    E = torch.tensor(E0, dtype=tdt, device=device)
    create_cholesky_factor_launch(L, E, dtype=tdt)()
    x1 = torch.tensor(b0, dtype=tdt, device=device)
    create_cholesky_solve_launch(L, E, x1, dtype=tdt)()
    L.tril_()
    x2 = torch.tensor(b0, dtype=tdt, device=device)
    create_cholesky_solve_launch(L, E, x2, dtype=tdt)()
    np.testing.assert_array_equal(x1.cpu().numpy(), x2.cpu().numpy())


class _SpinLimit:
    def __init__(self, polls):
        self.polls = polls

    def __enter__(self):
        from socu_amd import _lib

        _lib.load().socu_set_sweep_spin_limit(self.polls)

    def __exit__(self, *exc):
        from socu_amd import _lib

        _lib.load().socu_set_sweep_spin_limit(0)


@pytest.mark.parametrize("graph", [False, True])
def test_expired_wait_gives_nan_and_the_next_call_raises(graph, device):
    import torch
    from socu_amd import _lib
    from socu_amd.block_tridiag_solver import create_cholesky_factor_and_solve_launch

    n, N = 32, 64
    L0, E0, b0 = _stack(n, N, 1, 1, seed=8)
    L = torch.tensor(L0, device=device)
    E = torch.tensor(E0, device=device)
    x = torch.tensor(b0, device=device)
    with _SpinLimit(-1):
        run = create_cholesky_factor_and_solve_launch(L, E, x, use_cuda_graph=graph)
        run()
        torch.cuda.synchronize()
    assert np.isnan(x.cpu().numpy()).any()
    with pytest.raises(_lib.SweepExpiredError):
        run()
    # the word is cleared once reported; a launch made with the default bound
    # is unaffected
    L.copy_(torch.tensor(L0, device=device))
    E.copy_(torch.tensor(E0, device=device))
    x.copy_(torch.tensor(b0, device=device))
    ok = create_cholesky_factor_and_solve_launch(L, E, x, use_cuda_graph=graph)
    ok()
    assert ok.sweep_status() == (True, False)
    assert nd.rel_residual(L0, E0, x.cpu().numpy(), b0) < 1e-12


def test_expired_wait_reported_by_sweep_status(device):
    import torch
    from socu_amd.block_tridiag_solver import create_cholesky_solve_launch

    n, N = 16, 40
    L0, E0, b0 = _stack(n, N, 1, 1, seed=9)
    Lf, Ef = L0.copy(), E0.copy()
    nd.factor(Lf, Ef)
    L = torch.tensor(Lf, device=device)
    E = torch.tensor(Ef, device=device)
    x = torch.tensor(b0, device=device)
    with _SpinLimit(-1):
        run = create_cholesky_solve_launch(L, E, x)
        run()
    assert run.sweep_status() == (True, True)
    assert np.isnan(x.cpu().numpy()).any()
    # reported by sweep_status, so the next call runs (with the limit it was
    # enqueued under restored: the default)
    x.copy_(torch.tensor(b0, device=device))
    run()
    assert run.sweep_status() == (True, False)
    xr = b0.copy()
    nd.solve(Lf, Ef, xr)
    assert _rel(x.cpu().numpy(), xr) < 1e-11


def test_mixed_expired_wait_raises_next_call(device):
    import torch
    from socu_amd import _lib
    from socu_amd.mixed import create_cholesky_mixed_factor_and_solve_launch

    n, N = 32, 64
    L0, E0, b0 = _stack(n, N, 1, 1, seed=10)
    x = torch.tensor(b0, device=device)
    with _SpinLimit(-1):
        run = create_cholesky_mixed_factor_and_solve_launch(torch.tensor(L0, device=device),
                                                            torch.tensor(E0, device=device), x, max_iter=3,
                                                            use_cuda_graph=True)
        run()
        torch.cuda.synchronize()
    rr, _ = run.stats()
    assert not (rr[0] <= 1e-10)
    with pytest.raises(_lib.SweepExpiredError):
        run()


def test_direct_solve_entry_point_with_workspace(device):
    """socu_cholesky_solve through the C-ABI with a caller-provided workspace,
    on a foreign factor."""
    import torch
    from socu_amd import _lib

    lib = _lib.load()
    n, N, m, B = 40, 21, 3, 2
    L0, E0, b0 = _stack(n, N, m, B, seed=12)
    Lf, Ef = L0.copy(), E0.copy()
    nd.factor(Lf, Ef)
    xr = b0.copy()
    nd.solve(Lf, Ef, xr)
    L = torch.tensor(np.tril(Lf) + np.triu(np.ones_like(Lf), 1), device=device)
    E = torch.tensor(Ef, device=device)
    x = torch.tensor(b0, device=device)
    work = torch.empty(lib.socu_solve_workspace_size(B, N, n, _lib.DTYPE_F64), dtype=torch.uint8, device=device)
    st = ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
    _lib.check(lib.socu_cholesky_solve(L.data_ptr(), E.data_ptr(), x.data_ptr(), None, B, N, n, m, _lib.DTYPE_F64,
                                       work.data_ptr(), st), "solve")
    torch.cuda.synchronize()
    assert _rel(x.cpu().numpy(), xr) < 1e-11


@pytest.mark.parametrize("n,N,m,dtype", [(32, 37, 1, "float64"), (128, 24, 2, "float64"), (64, 50, 1, "float32"),
                                         (10, 20, 3, "float64")])
def test_factor_ignores_the_fill_in_region_on_entry(n, N, m, dtype, device):
    """The reference needs E[N-1:] zeroed before each factor (it accumulates
    F -= R Lc there); this factor writes each fill-in block once, so noise in
    that region on entry must give the oracle's factor and solution."""
    import torch
    from socu_amd.block_tridiag_solver import create_cholesky_factor_and_solve_launch

    tdt = getattr(torch, dtype)
    L0, E0, b0 = _stack(n, N, m, 2, seed=31)
    Lr, Er, xr = L0.copy(), E0.copy(), b0.copy()
    nd.factor_and_solve(Lr, Er, xr)
    En = E0.copy()
    En[:, N - 1:] = np.random.RandomState(6).randn(*En[:, N - 1:].shape) * 1e6
    L = torch.tensor(L0, dtype=tdt, device=device)
    E = torch.tensor(En, dtype=tdt, device=device)
    x = torch.tensor(b0, dtype=tdt, device=device)
    create_cholesky_factor_and_solve_launch(L, E, x, dtype=tdt)()
    tol = 1e-11 if dtype == "float64" else 2e-4
    assert _rel(np.tril(L.cpu().numpy()), np.tril(Lr)) < tol
    assert _rel(E.cpu().numpy(), Er) < tol
    assert _rel(x.cpu().numpy(), xr) < tol


def test_batched_factor_ignores_the_fill_in_region_on_entry(device):
    import torch
    from socu_amd.batched import create_batched_cholesky_launch

    systems, refs = [], []
    rs = np.random.RandomState(7)
    for k, (n, N) in enumerate([(8, 33), (16, 20), (24, 47), (32, 64), (12, 5)]):
        L0, E0, b0, _, _ = nd.readme_problem(n, N, 1, alpha=10.0, seed=50 + k)
        Lr, Er, xr = L0.copy(), E0.copy(), b0.copy()
        nd.factor_and_solve(Lr, Er, xr)
        En = E0.copy()
        En[N - 1:] = rs.randn(*En[N - 1:].shape) * 1e6
        systems.append((torch.tensor(L0, device=device), torch.tensor(En, device=device),
                        torch.tensor(b0, device=device)))
        refs.append((Lr, Er, xr))
    create_batched_cholesky_launch([s[0] for s in systems], [s[1] for s in systems], [s[2] for s in systems],
                                   op="factor_and_solve", tune=False)()
    torch.cuda.synchronize()
    for (L, E, x), (Lr, Er, xr) in zip(systems, refs):
        assert _rel(np.tril(L.cpu().numpy()), np.tril(Lr)) < 1e-11
        assert _rel(E.cpu().numpy(), Er) < 1e-11
        assert _rel(x.cpu().numpy(), xr) < 1e-11


@pytest.mark.parametrize("n,N,m,dtype", [(32, 40, 1, "float64"), (96, 17, 3, "float32"), (200, 9, 2, "float64")])
def test_solve_with_stored_inverses_equals_solve(n, N, m, dtype, device):
    """socu_cholesky_solve_with_inverses on this library's own factor gives
    socu_cholesky_solve's x to rounding: the inverses it reads in place (made
    by the factor) and the ones the solve forms from tril(L) come from
    different kernels, so they may differ in the last bits."""
    import torch
    from socu_amd import _lib

    tdt = getattr(torch, dtype)
    code = _lib.DTYPE_F64 if dtype == "float64" else _lib.DTYPE_F32
    lib = _lib.load()
    B = 2
    L0, E0, b0 = _stack(n, N, m, B, seed=4)
    L = torch.tensor(L0, dtype=tdt, device=device)
    E = torch.tensor(E0, dtype=tdt, device=device)
    st = ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
    _lib.check(lib.socu_cholesky_factor(L.data_ptr(), E.data_ptr(), None, B, N, n, code, st), "factor")
    x1 = torch.tensor(b0, dtype=tdt, device=device)
    x2 = x1.clone()
    work = torch.empty(lib.socu_solve_workspace_size(B, N, n, code), dtype=torch.uint8, device=device)
    _lib.check(lib.socu_cholesky_solve(L.data_ptr(), E.data_ptr(), x1.data_ptr(), None, B, N, n, m, code,
                                       work.data_ptr(), st), "solve")
    _lib.check(lib.socu_cholesky_solve_with_inverses(L.data_ptr(), E.data_ptr(), x2.data_ptr(), None, B, N, n, m,
                                                     code, st), "solve_with_inverses")
    torch.cuda.synchronize(device)
    a, b = x1.cpu().numpy(), x2.cpu().numpy()
    assert np.linalg.norm(b - a) / np.linalg.norm(a) < (1e-13 if dtype == "float64" else 1e-5)


def _capi_codes(dtype):
    from socu_amd import _lib
    return _lib.load(), (_lib.DTYPE_F64 if dtype == "float64" else _lib.DTYPE_F32)


@pytest.mark.parametrize("n,N,m,dtype", [(128, 24, 1, "float64"), (40, 33, 2, "float64"), (64, 50, 1, "float32"),
                                         (32, 1, 1, "float64"), (8, 2, 3, "float32"), (160, 9, 2, "float64"),
                                         (32, 45, 40, "float32")])
def test_out_of_place_factor_equals_in_place(n, N, m, dtype, device):
    """socu_cholesky_factor_from / _and_solve_from: bit-identical to the
    in-place entry points on copies of the inputs (the same kernels read the
    first level's blocks from the sources), the sources untouched, masked
    systems' outputs not written.  n = 160 takes the copy-then-factor path."""
    import torch

    lib, code = _capi_codes(dtype)
    tdt = getattr(torch, dtype)
    B = 3
    L0, E0, b0 = _stack(n, N, m, B, seed=71, alpha=32.0 if n > 128 else 16.0)
    st = torch.cuda.current_stream(device).cuda_stream
    mask = torch.tensor([1, 0, 1], dtype=torch.uint8, device=device)
    for with_solve in (False, True):
        Ls = torch.tensor(L0, dtype=tdt, device=device)
        Es = torch.tensor(E0, dtype=tdt, device=device)
        Li, Ei, xi = Ls.clone(), Es.clone(), torch.tensor(b0, dtype=tdt, device=device)
        Lo = torch.full_like(Ls, 7.0)
        Eo = torch.full_like(Es, 7.0)
        xo = xi.clone()
        if with_solve:
            assert lib.socu_cholesky_factor_and_solve(Li.data_ptr(), Ei.data_ptr(), xi.data_ptr(), mask.data_ptr(), B,
                                                      N, n, m, code, st) == 0
            assert lib.socu_cholesky_factor_and_solve_from(Ls.data_ptr(), Es.data_ptr(), Lo.data_ptr(), Eo.data_ptr(),
                                                           xo.data_ptr(), mask.data_ptr(), B, N, n, m, code, st) == 0
        else:
            assert lib.socu_cholesky_factor(Li.data_ptr(), Ei.data_ptr(), mask.data_ptr(), B, N, n, code, st) == 0
            assert lib.socu_cholesky_factor_from(Ls.data_ptr(), Es.data_ptr(), Lo.data_ptr(), Eo.data_ptr(),
                                                 mask.data_ptr(), B, N, n, code, st) == 0
        torch.cuda.synchronize()
        np.testing.assert_array_equal(Ls.cpu().numpy(), np.asarray(L0, dtype=dtype))
        np.testing.assert_array_equal(Es.cpu().numpy(), np.asarray(E0, dtype=dtype))
        for b in (0, 2):
            np.testing.assert_array_equal(Lo[b].cpu().numpy(), Li[b].cpu().numpy())
            np.testing.assert_array_equal(Eo[b].cpu().numpy(), Ei[b].cpu().numpy())
            np.testing.assert_array_equal(xo[b].cpu().numpy(), xi[b].cpu().numpy())
        if n <= 128:  # masked system: outputs not written
            assert bool((Lo[1] == 7.0).all()) and bool((Eo[1] == 7.0).all())
        np.testing.assert_array_equal(xo[1].cpu().numpy(), np.asarray(b0[1], dtype=dtype))
