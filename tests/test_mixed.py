"""Mixed-precision factor_and_solve with FP64 refinement (an extension; the
reference has no such path).  Checks: the FP64 relative residual reaches the
requested tolerance, the solution matches the oracle's FP64 solution to the
accuracy the residual implies, Psi (L, E) is untouched, masked entries are
untouched, and graph replays agree with eager calls."""
import numpy as np
import pytest

from oracle import nd

pytestmark = pytest.mark.gpu


def _problem(n, N, m, alpha, seed, B=1):
    Ls, Es, bs = [], [], []
    for k in range(B):
        L0, E0, b0, _, _ = nd.readme_problem(n, N, m, alpha=alpha, seed=seed + k)
        Ls.append(L0)
        Es.append(E0)
        bs.append(b0)
    return np.stack(Ls), np.stack(Es), np.stack(bs)


@pytest.mark.parametrize("n,N,m,alpha", [(32, 20, 1, 10.0), (64, 33, 4, 10.0), (128, 16, 1, 16.0), (36, 7, 3, 10.0),
                                         (128, 64, 8, 16.0), (10, 100, 2, 10.0),
                                         # one right-hand side (two-pass residual) at ragged n, one block, two blocks
                                         (36, 9, 1, 10.0), (10, 33, 1, 10.0), (100, 5, 1, 16.0), (16, 1, 1, 10.0),
                                         (24, 2, 1, 10.0)])
@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("precision", ["float32", "float32_f32_pivots"])
def test_mixed_reaches_fp64_residual(n, N, m, alpha, graph, precision, device):
    import torch
    from socu_amd.mixed import create_cholesky_mixed_factor_and_solve_launch

    L0, E0, b0 = _problem(n, N, m, alpha, seed=11, B=2)
    L = torch.tensor(L0, device=device)
    E = torch.tensor(E0, device=device)
    x = torch.tensor(b0, device=device)
    run = create_cholesky_mixed_factor_and_solve_launch(L, E, x, precision=precision, tol=1e-10, max_iter=8,
                                                        use_cuda_graph=graph)
    for rep in range(2):  # second call: graph replay / re-run on fresh b
        x.copy_(torch.tensor(b0, device=device))
        run()
        rr, its = run.stats()
        xg = x.cpu().numpy()
        res = nd.rel_residual(L0, E0, xg, b0)
        assert np.all(res <= 1e-10), (res, its)
        # the FP64 residual itself carries ~1e-16 * ||Psi|| ||x|| / ||b|| of
        # rounding (summed over n blocks and m columns); agreement to 1e-14
        # absolute is four orders below the 1e-10 tolerance being checked
        np.testing.assert_allclose(res, rr, rtol=0, atol=1e-14)
        assert np.all(its >= 1) and np.all(its <= 8)
        Lr, Er, xr = L0.copy(), E0.copy(), b0.copy()
        nd.factor_and_solve(Lr, Er, xr)
        assert np.linalg.norm(xg - xr) / np.linalg.norm(xr) < 1e-8
    # Psi untouched
    np.testing.assert_array_equal(L.cpu().numpy(), L0)
    np.testing.assert_array_equal(E.cpu().numpy(), E0)


def test_mixed_mask(device):
    import torch
    from socu_amd.mixed import create_cholesky_mixed_factor_and_solve_launch

    L0, E0, b0 = _problem(32, 9, 2, 10.0, seed=3, B=4)
    L = torch.tensor(L0, device=device)
    E = torch.tensor(E0, device=device)
    x = torch.tensor(b0, device=device)
    mask = torch.tensor([True, False, True, False], device=device)
    run = create_cholesky_mixed_factor_and_solve_launch(L, E, x, mask, use_cuda_graph=True)
    run()
    xg = x.cpu().numpy()
    res = nd.rel_residual(L0, E0, xg, b0)
    assert res[0] <= 1e-10 and res[2] <= 1e-10
    np.testing.assert_array_equal(xg[1], b0[1])
    np.testing.assert_array_equal(xg[3], b0[3])


def test_mixed_needs_fewer_than_max_iter_for_well_conditioned(device):
    import torch
    from socu_amd.mixed import create_cholesky_mixed_factor_and_solve_launch

    L0, E0, b0 = _problem(128, 32, 1, 16.0, seed=0)
    x = torch.tensor(b0, device=device)
    run = create_cholesky_mixed_factor_and_solve_launch(torch.tensor(L0, device=device),
                                                        torch.tensor(E0, device=device), x, max_iter=10)
    run()
    rr, its = run.stats()
    assert rr[0] <= 1e-10 and its[0] <= 3, (rr, its)


@pytest.mark.parametrize("B", [1, 3])
@pytest.mark.parametrize("graph", [False, True])
def test_mixed_iterations_after_convergence_leave_x_unchanged(B, graph, device):
    """Refinement steps past convergence (every system converged: the
    persistent sweep ends at once) must not touch x: the result of a solve
    capped at the number of steps it needs is reproduced bit for bit with a
    higher cap, on repeated calls."""
    import torch
    from socu_amd.mixed import create_cholesky_mixed_factor_and_solve_launch

    L0, E0, b0 = _problem(128, 64, 1, 16.0, seed=3, B=B)
    Lt, Et = torch.tensor(L0, device=device), torch.tensor(E0, device=device)
    x_need = torch.tensor(b0, device=device)
    probe = create_cholesky_mixed_factor_and_solve_launch(Lt, Et, x_need, max_iter=10, use_cuda_graph=graph)
    probe()
    rr, its = probe.stats()
    need = int(max(its))
    assert need < 10 and max(rr) <= 1e-10, (rr, its)
    for cap in (need, need + 3):
        x = torch.tensor(b0, device=device)
        run = create_cholesky_mixed_factor_and_solve_launch(Lt, Et, x, max_iter=cap, use_cuda_graph=graph)
        for _ in range(2):
            x.copy_(torch.tensor(b0, device=device))
            run()
            torch.cuda.synchronize()
            assert torch.equal(x, x_need), cap
            assert [int(v) for v in run.stats()[1]] == [int(v) for v in its]


def test_mixed_non_spd_stops_without_hanging(device):
    import torch
    from socu_amd.mixed import create_cholesky_mixed_factor_and_solve_launch

    L0, E0, b0 = _problem(8, 5, 1, 10.0, seed=1)
    L0[0, 2] = -np.eye(8)
    x = torch.tensor(b0, device=device)
    run = create_cholesky_mixed_factor_and_solve_launch(torch.tensor(L0, device=device),
                                                        torch.tensor(E0, device=device), x, max_iter=4)
    run()
    rr, its = run.stats()
    assert not (rr[0] <= 1e-10)
    assert its[0] <= 4
