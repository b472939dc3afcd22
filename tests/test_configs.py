"""The five benchmark configurations at their full sizes, against the oracle.

C1  FP64 factor then solve, n=32 N=100 m=1, README generator seed 0: the
    dense-matrix residual check of the README (< 1e-8) and x vs the oracle.
C2  FP64 factor_and_solve, n=64 N=1024 m=16: tril(L) and every E index
    (fill-in included, zero pattern exact) vs the oracle to 1e-11 relative,
    rel-resid <= 1e-12.
C3  mixed FP32 factor + FP64 refinement, n=128 N=512, m in {1, 8}, both pivot
    precisions, alpha=16 (well conditioned): rel-resid <= 1e-10 and x vs the
    FP64 oracle; alpha in {10, 2} (numerically singular at n=128): the GPU
    factor breaks down where the oracle's does, and the mixed solve stops
    unconverged without hanging.
C5  differentiable solve, B=256 n=32 N=128 m=4, README generator seed 2:
    central finite differences on 64 random entries.  At h=1e-6 the
    difference quotient carries the FP64 rounding of the loss (~2e-15 on a
    loss of ~1, so ~1e-9 absolute on gradients of 1e-3..5e-2: 2.5e-8 relative
    measured with the numpy oracle's own dense solve on this problem), so h=1e-6
    is held to 1e-7; at h=1e-4 (truncation and rounding both ~3e-10 relative
    on the oracle) the gradients are held to 1e-8.

The large oracle solves use the oracle's C restatement (oracle/nd_c.py,
pinned against the numpy oracle in tests/test_oracle.py).
"""
import numpy as np
import pytest

from oracle import nd, nd_c

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300)


def _c_factor_and_solve(L0, E0, b0):
    L, E, x = L0[None].copy(), E0[None].copy(), b0[None].copy()
    nd_c.factor_and_solve(L, E, x)
    return L[0], E[0], x[0]


def test_c1_readme_generator_factor_then_solve(device):
    import torch
    from socu_amd.block_tridiag_solver import create_cholesky_factor_launch, create_cholesky_solve_launch

    n, N, m = 32, 100, 1
    L0, E0, b0, _, _ = nd.readme_problem(n, N, m, alpha=10.0, seed=0)
    L = torch.tensor(L0, device=device)
    E = torch.tensor(E0, device=device)
    x = torch.tensor(b0, device=device)
    factor = create_cholesky_factor_launch(L, E)
    solve = create_cholesky_solve_launch(L, E, x)
    factor()
    solve()
    xg = x.cpu().numpy()
    A = nd.dense_from_blocks(L0, E0)
    assert np.linalg.norm(A @ xg.reshape(-1) - b0.reshape(-1)) < 1e-8  # README.md:146-156
    xr = b0.copy()
    Lr, Er = L0.copy(), E0.copy()
    nd.factor_and_solve(Lr, Er, xr)
    assert _rel(xg, xr) < 1e-12


@pytest.fixture(scope="module")
def c2_problem():
    n, N, m = 64, 1024, 16
    L0, E0, b0, _, _ = nd.readme_problem(n, N, m, alpha=10.0, seed=0)
    Lr, Er, xr = _c_factor_and_solve(L0, E0, b0)
    return L0, E0, b0, Lr, Er, xr


@pytest.mark.parametrize("graph", [False, True])
def test_c2_fp64_factor_and_solve_n64_N1024_m16(c2_problem, graph, device):
    import torch
    from socu_amd.block_tridiag_solver import create_cholesky_factor_and_solve_launch

    L0, E0, b0, Lr, Er, xr = c2_problem
    L = torch.tensor(L0, device=device)
    E = torch.tensor(E0, device=device)
    x = torch.tensor(b0, device=device)
    create_cholesky_factor_and_solve_launch(L, E, x, use_cuda_graph=graph)()
    Lg, Eg, xg = L.cpu().numpy(), E.cpu().numpy(), x.cpu().numpy()
    assert _rel(np.tril(Lg), np.tril(Lr)) < 1e-11
    worst = max(_rel(np.tril(Lg[i]), np.tril(Lr[i])) for i in range(Lr.shape[0]))
    assert worst < 1e-11, worst
    # every E index: the same blocks are zero (fill-in indexing), the others
    # agree block by block
    zr = np.all(Er == 0, axis=(1, 2))
    zg = np.all(Eg == 0, axis=(1, 2))
    np.testing.assert_array_equal(zg, zr)
    worst = max(_rel(Eg[k], Er[k]) for k in np.where(~zr)[0])
    assert worst < 1e-11, worst
    assert _rel(xg, xr) < 1e-11
    # rel-resid <= 1e-12, or no worse than the FP64 oracle's own on this
    # problem (its factor_and_solve reaches 1.39e-12 here: the bound is at the
    # problem's FP64 limit)
    res = nd.rel_residual(L0, E0, xg, b0)
    res_oracle = nd.rel_residual(L0, E0, xr, b0)
    print(f"C2 rel-resid: GPU {res:.3e}, oracle {res_oracle:.3e}")
    assert res <= max(1e-12, res_oracle)


@pytest.fixture(scope="module")
def c3_problems():
    out = {}
    for m in (1, 8):
        L0, E0, b0, _, _ = nd.readme_problem(128, 512, m, alpha=16.0, seed=0)
        _, _, xr = _c_factor_and_solve(L0, E0, b0)
        out[m] = (L0, E0, b0, xr)
    return out


@pytest.mark.parametrize("m", [1, 8])
@pytest.mark.parametrize("precision", ["float32", "float32_f32_pivots"])
def test_c3_mixed_n128_N512(c3_problems, m, precision, device):
    import torch
    from socu_amd.mixed import create_cholesky_mixed_factor_and_solve_launch

    L0, E0, b0, xr = c3_problems[m]
    x = torch.tensor(b0, device=device)
    run = create_cholesky_mixed_factor_and_solve_launch(torch.tensor(L0, device=device),
                                                        torch.tensor(E0, device=device), x, precision=precision,
                                                        tol=1e-10, max_iter=10, use_cuda_graph=True)
    for _ in range(2):
        x.copy_(torch.tensor(b0, device=device))
        run()
        rr, its = run.stats()
        xg = x.cpu().numpy()
        res = nd.rel_residual(L0, E0, xg, b0)
        assert res <= 1e-10, (res, its)
        assert 1 <= its[0] <= 4, its
        # kappa(Psi) is ~1e3 at alpha = 16: x agrees with the FP64 solution to
        # about kappa * 1e-10
        assert _rel(xg, xr) < 1e-7


@pytest.mark.parametrize("alpha", [10.0, 2.0])
def test_c3_literal_alpha_breaks_down_like_the_oracle(alpha, device):
    """At n=128 the README generator is numerically singular for alpha=10
    and alpha=2: the ND factor meets a non-positive pivot.  The GPU's FP64
    factor breaks down at the same ND level as the oracle's (the first level
    with a non-finite diagonal block), the solution is non-finite as the
    oracle's is, and the mixed solve reports non-convergence."""
    import torch
    from socu_amd.block_tridiag_solver import create_cholesky_factor_and_solve_launch
    from socu_amd.mixed import create_cholesky_mixed_factor_and_solve_launch

    n, N = 128, 512
    L0, E0, b0, _, _ = nd.readme_problem(n, N, 1, alpha=alpha, seed=0)
    Lr, _, xr = _c_factor_and_solve(L0, E0, b0)

    def first_bad_level(Lf):
        bad = ~np.isfinite(np.diagonal(Lf, axis1=1, axis2=2)).all(1)
        for lev in nd.level_schedule(N):
            if bad[lev["nodes"]].any():
                return lev["l"]
        return None

    ref_level = first_bad_level(Lr)
    assert ref_level is not None and not np.isfinite(xr).all()
    L = torch.tensor(L0, device=device)
    E = torch.tensor(E0, device=device)
    x = torch.tensor(b0, device=device)
    create_cholesky_factor_and_solve_launch(L, E, x)()
    gpu_level = first_bad_level(L.cpu().numpy())
    if gpu_level != ref_level:
        # Only where a level-0 block is itself numerically singular may the
        # breakdown come earlier: whether a pivot of such a block rounds to a
        # negative value depends on the summation order (at alpha=2, L_0 =
        # D_0 D_0^T has lambda_min = -2.9e-14 against lambda_max = 344 in FP64,
        # and the oracle's order happens to keep its pivots positive).
        assert gpu_level == 0 and ref_level > 0
        lam = [np.linalg.eigvalsh(L0[i]) for i in nd.level_schedule(N)[0]["nodes"]]
        assert min(v[0] / v[-1] for v in lam) < 1e-12
    assert not np.isfinite(x.cpu().numpy()).all()
    xm = torch.tensor(b0, device=device)
    run = create_cholesky_mixed_factor_and_solve_launch(L.new_tensor(L0), E.new_tensor(E0), xm, max_iter=4)
    run()
    rr, its = run.stats()
    assert not (rr[0] <= 1e-10) and its[0] <= 4


def test_c5_finite_differences_at_spec(device):
    """B=256, n=32, N=128, m=4, README generator seed 2 (one stream, systems
    drawn in turn, then w).  64 entries, each in its own system, are
    perturbed at once by +-h; the per-system loss sum(w_b * x_b)
    difference is each entry's central difference (systems are independent,
    and the kernels compute each system identically whatever the others hold).
    Entries: lower triangles of L blocks, coupling blocks E[:N-1], and b."""
    import torch
    from socu_amd.autograd import block_tridiag_solve

    B, n, N, m = 256, 32, 128, 4
    L0, E0, b0, w0 = nd.readme_batch(B, n, N, m, alpha=10.0, seed=2, with_weights=True)
    L = torch.tensor(L0, device=device, requires_grad=True)
    E = torch.tensor(E0, device=device, requires_grad=True)
    b = torch.tensor(b0, device=device, requires_grad=True)
    w = torch.tensor(w0, device=device)
    (block_tridiag_solve(L, E, b) * w).sum().backward()
    gL, gE, gb = L.grad.cpu().numpy(), E.grad.cpu().numpy(), b.grad.cpu().numpy()

    rs = np.random.RandomState(64)
    systems = rs.choice(B, 64, replace=False)
    picks = []
    for k, sys_ in enumerate(systems):
        kind = k % 3
        if kind == 0:
            i = rs.randint(N)
            r = rs.randint(n)
            c = rs.randint(r + 1)
            picks.append(("L", (sys_, i, r, c)))
        elif kind == 1:
            picks.append(("E", (sys_, rs.randint(N - 1), rs.randint(n), rs.randint(n))))
        else:
            picks.append(("b", (sys_, rs.randint(N), rs.randint(n), rs.randint(m))))
    def losses(sign, h):
        Lp, Ep, bp = L0.copy(), E0.copy(), b0.copy()
        for kind, idx in picks:
            {"L": Lp, "E": Ep, "b": bp}[kind][idx] += sign * h
        with torch.no_grad():
            x = block_tridiag_solve(torch.tensor(Lp, device=device), torch.tensor(Ep, device=device),
                                    torch.tensor(bp, device=device))
            return (x * w).sum(dim=(1, 2, 3)).cpu().numpy()

    an = np.array([{"L": gL, "E": gE, "b": gb}[kind][idx] for kind, idx in picks])
    for h, tol in ((1e-6, 1e-7), (1e-4, 1e-8)):
        lp, lm = losses(+1.0, h), losses(-1.0, h)
        fd = np.array([(lp[idx[0]] - lm[idx[0]]) / (2 * h) for _, idx in picks])
        err = np.linalg.norm(fd - an) / np.linalg.norm(an)
        print(f"C5 central differences h={h:g}: rel-err {err:.2e} over 64 entries")
        assert err <= tol, (h, err, np.max(np.abs(fd - an)))
