"""Batched launches over uneven systems, checked per system against the oracle.

Each system of a batch is an independent instance of the reference's problem,
so the oracle's factor / solve of that system alone is the expected result
(FP64: relative error <= 1e-12; FP32: <= 1e-4).  The full 4096-system MPC
batch is checked through the size-independent relative residual (<= 1e-10 for
every system) plus an oracle comparison on a sample.
"""
import numpy as np
import pytest

from oracle import nd

pytestmark = pytest.mark.gpu

# (N, n, m): horizons from one block to several levels, block sizes on and
# off the tile classes (8, 16, 32), several right-hand-side widths.
SHAPES = [(1, 1, 1), (1, 8, 2), (2, 3, 1), (3, 5, 2), (5, 8, 1), (8, 9, 3), (17, 12, 1), (31, 16, 4), (64, 17, 1),
          (100, 24, 2), (33, 31, 1), (45, 32, 8), (7, 32, 32), (12, 4, 20), (200, 8, 1), (129, 32, 1)]


def _make(shapes, seed=5):
    out = []
    for k, (N, n, m) in enumerate(shapes):
        L, E, b, _, _ = nd.readme_problem(n, N, m, alpha=10.0, seed=seed + k)
        out.append((L, E, b))
    return out


def _dev(systems, device, dtype):
    import torch

    Ls = [torch.tensor(L, device=device, dtype=dtype) for L, _, _ in systems]
    Es = [torch.tensor(E, device=device, dtype=dtype) for _, E, _ in systems]
    xs = [torch.tensor(b, device=device, dtype=dtype) for _, _, b in systems]
    return Ls, Es, xs


def _oracle(L, E, b, op="factor_and_solve"):
    L, E, x = L.copy(), E.copy(), b.copy()
    if op in ("factor", "factor_and_solve"):
        nd.factor(L, E)
    if op == "factor_and_solve":
        nd.solve(L, E, x)
    return L, E, x


def _rel(a, b):
    d = np.linalg.norm(a - b)
    s = np.linalg.norm(b)
    return d / s if s > 0 else d


@pytest.mark.parametrize("tune,graph", [(False, False), (True, False), (True, True)])
def test_uneven_factor_and_solve_matches_oracle(tune, graph, device):
    import torch
    from socu_amd.batched import create_batched_cholesky_launch

    systems = _make(SHAPES)
    Ls, Es, xs = _dev(systems, device, torch.float64)
    run = create_batched_cholesky_launch(Ls, Es, xs, op="factor_and_solve", tune=tune, use_cuda_graph=graph)
    run()
    torch.cuda.synchronize()
    for k, (L0, E0, b0) in enumerate(systems):
        Lr, Er, xr = _oracle(L0, E0, b0)
        assert _rel(np.tril(Ls[k].cpu().numpy()), np.tril(Lr)) < 1e-12, (k, SHAPES[k])
        assert _rel(Es[k].cpu().numpy(), Er) < 1e-12, (k, SHAPES[k])
        assert _rel(xs[k].cpu().numpy(), xr) < 1e-12, (k, SHAPES[k])
        assert nd.rel_residual(L0[None], E0[None], xs[k].cpu().numpy()[None], b0[None])[0] < 1e-12
    cfg = run.config()
    assert sum(v["systems"] for v in cfg.values()) == len(SHAPES)
    assert all(v["tile"] in (8, 16, 32) and v["tile"] >= bound and v["waves"] in (1, 2, 4) for bound, v in cfg.items())


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("waves", [1, 2, 4])
def test_every_workgroup_size_matches_oracle(waves, dtype, device, monkeypatch):
    """Untuned plans with 1, 2 and 4 wavefronts per workgroup (every bucket),
    so that each variant of the kernels runs whatever the tuner would pick.
    Workgroups whose LDS would not fit fall back to fewer wavefronts (FP64
    tiles of 32 with four tasks per workgroup leave room for m <= 7 only):
    the full shape list, then the same shapes with m <= 4, where every bucket
    takes the requested width."""
    import torch
    from socu_amd.batched import create_batched_cholesky_launch

    monkeypatch.setenv("SOCU_BATCHED_WAVES", str(waves))
    tdt = getattr(torch, dtype)
    tol = 1e-12 if dtype == "float64" else 2e-4
    for shapes in (SHAPES, [(N, n, min(m, 4)) for N, n, m in SHAPES]):
        systems = _make(shapes)
        Ls, Es, xs = _dev(systems, device, tdt)
        run = create_batched_cholesky_launch(Ls, Es, xs, op="factor_and_solve", tune=False, dtype=tdt)
        run()
        torch.cuda.synchronize()
        cfg = run.config()
        assert all(v["waves"] <= waves for v in cfg.values())
        if max(m for _, _, m in shapes) <= 4:
            assert all(v["waves"] == waves for v in cfg.values()), cfg
        for k, (L0, E0, b0) in enumerate(systems):
            Lr, Er, xr = _oracle(L0, E0, b0)
            assert _rel(np.tril(Ls[k].double().cpu().numpy()), np.tril(Lr)) < tol, (k, shapes[k])
            assert _rel(Es[k].double().cpu().numpy(), Er) < tol, (k, shapes[k])
            assert _rel(xs[k].double().cpu().numpy(), xr) < tol, (k, shapes[k])


def test_graph_replays_match_eager(device):
    import torch
    from socu_amd.batched import create_batched_cholesky_launch

    systems = _make(SHAPES[:10], seed=21)
    La, Ea, xa = _dev(systems, device, torch.float64)
    create_batched_cholesky_launch(La, Ea, xa, tune=False)()
    Lg, Eg, xg = _dev(systems, device, torch.float64)
    run = create_batched_cholesky_launch(Lg, Eg, xg, tune=False, use_cuda_graph=True)
    pristine = _dev(systems, device, torch.float64)
    for rep in range(3):
        for dst, src in zip(Lg + Eg + xg, pristine[0] + pristine[1] + pristine[2]):
            dst.copy_(src)
        torch.cuda.synchronize()
        run()
        torch.cuda.synchronize()
        for a, b in zip(Lg + Eg + xg, La + Ea + xa):
            assert torch.equal(a, b), rep


def test_split_operations_match_oracle(device):
    """factor, then solve; factor, then forward and backward substitution."""
    import torch
    from socu_amd.batched import create_batched_cholesky_launch

    systems = _make(SHAPES, seed=40)
    Ls, Es, xs = _dev(systems, device, torch.float64)
    create_batched_cholesky_launch(Ls, Es, op="factor", tune=False)()
    x2 = [x.clone() for x in xs]
    create_batched_cholesky_launch(Ls, Es, xs, op="solve", tune=False)()
    create_batched_cholesky_launch(Ls, Es, x2, op="forward_substitution", tune=False)()
    create_batched_cholesky_launch(Ls, Es, x2, op="backward_substitution", tune=False)()
    torch.cuda.synchronize()
    for k, (L0, E0, b0) in enumerate(systems):
        Lr, Er, xr = _oracle(L0, E0, b0)
        assert _rel(np.tril(Ls[k].cpu().numpy()), np.tril(Lr)) < 1e-12
        assert _rel(Es[k].cpu().numpy(), Er) < 1e-12
        assert _rel(xs[k].cpu().numpy(), xr) < 1e-12, (k, SHAPES[k])
        assert torch.equal(xs[k], x2[k]) or _rel(x2[k].cpu().numpy(), xr) < 1e-12


def test_forward_only_matches_oracle(device):
    import torch
    from socu_amd.batched import create_batched_cholesky_launch

    systems = _make(SHAPES[:9], seed=60)
    Ls, Es, xs = _dev(systems, device, torch.float64)
    create_batched_cholesky_launch(Ls, Es, op="factor", tune=False)()
    create_batched_cholesky_launch(Ls, Es, xs, op="forward_substitution", tune=False)()
    torch.cuda.synchronize()
    for k, (L0, E0, b0) in enumerate(systems):
        Lr, Er, y = L0.copy(), E0.copy(), b0.copy()
        nd.factor(Lr, Er)
        nd.forward(Lr, Er, y)
        assert _rel(xs[k].cpu().numpy(), y) < 1e-12, (k, SHAPES[k])


def test_mask_leaves_systems_untouched(device):
    import torch
    from socu_amd.batched import create_batched_cholesky_launch

    systems = _make(SHAPES[:8], seed=70)
    Ls, Es, xs = _dev(systems, device, torch.float64)
    mask = torch.tensor([k % 3 != 1 for k in range(8)], device=device)
    run = create_batched_cholesky_launch(Ls, Es, xs, mask=mask, tune=False, use_cuda_graph=True)
    run()
    torch.cuda.synchronize()
    for k, (L0, E0, b0) in enumerate(systems):
        if k % 3 == 1:
            np.testing.assert_array_equal(Ls[k].cpu().numpy(), L0)
            np.testing.assert_array_equal(Es[k].cpu().numpy(), E0)
            np.testing.assert_array_equal(xs[k].cpu().numpy(), b0)
        else:
            assert _rel(xs[k].cpu().numpy(), _oracle(L0, E0, b0)[2]) < 1e-12
    # flip the mask in place: the captured graph reads it at execution time
    pristine = _dev(systems, device, torch.float64)
    for dst, src in zip(Ls + Es + xs, pristine[0] + pristine[1] + pristine[2]):
        dst.copy_(src)
    mask.copy_(~mask)
    run()
    torch.cuda.synchronize()
    for k, (L0, E0, b0) in enumerate(systems):
        if k % 3 != 1:
            np.testing.assert_array_equal(xs[k].cpu().numpy(), b0)
        else:
            assert _rel(xs[k].cpu().numpy(), _oracle(L0, E0, b0)[2]) < 1e-12


def test_float32_batch(device):
    import torch
    from socu_amd.batched import create_batched_cholesky_launch

    systems = _make(SHAPES, seed=80)
    Ls, Es, xs = _dev(systems, device, torch.float32)
    create_batched_cholesky_launch(Ls, Es, xs, dtype=torch.float32)()
    torch.cuda.synchronize()
    for k, (L0, E0, b0) in enumerate(systems):
        assert _rel(xs[k].cpu().numpy().astype(np.float64), _oracle(L0, E0, b0)[2]) < 1e-4, (k, SHAPES[k])


def test_batched_factor_feeds_the_regular_solve(device):
    """The batched factor stores what the regular launches expect, and back."""
    import torch
    from socu_amd.batched import create_batched_cholesky_launch
    from socu_amd.block_tridiag_solver import create_cholesky_factor_launch, create_cholesky_solve_launch

    L0, E0, b0, _, _ = nd.readme_problem(32, 40, 3, alpha=10.0, seed=90)
    xr = _oracle(L0, E0, b0)[2]
    L, E, x = (torch.tensor(a, device=device) for a in (L0, E0, b0))
    create_batched_cholesky_launch([L], [E], op="factor", tune=False)()
    create_cholesky_solve_launch(L, E, x)()
    torch.cuda.synchronize()
    assert _rel(x.cpu().numpy(), xr) < 1e-12
    L, E, x = (torch.tensor(a, device=device) for a in (L0, E0, b0))
    create_cholesky_factor_launch(L, E)()
    create_batched_cholesky_launch([L], [E], [x], op="solve", tune=False)()
    torch.cuda.synchronize()
    assert _rel(x.cpu().numpy(), xr) < 1e-12


def test_deterministic(device):
    import torch
    from socu_amd.batched import create_batched_cholesky_launch

    systems = _make(SHAPES, seed=95)
    outs = []
    for _ in range(2):
        Ls, Es, xs = _dev(systems, device, torch.float64)
        create_batched_cholesky_launch(Ls, Es, xs, tune=False)()
        torch.cuda.synchronize()
        outs.append([t.cpu().numpy() for t in Ls + Es + xs])
    for a, b in zip(*outs):
        np.testing.assert_array_equal(a, b)


def test_mpc_batch_4096(device):
    """The full uneven MPC batch: 4096 systems, N in [10, 200], n in
    {8, 12, 16, 24, 32}, m = 1, FP64, tuned and graph-captured."""
    import torch
    from socu_amd import problems
    from socu_amd.batched import create_batched_cholesky_launch

    batch = problems.mpc_batch(4096, seed=1)
    Ls, Es, xs = _dev(batch, device, torch.float64)
    run = create_batched_cholesky_launch(Ls, Es, xs, tune=True, use_cuda_graph=True)
    run()
    torch.cuda.synchronize()
    worst = 0.0
    for k, (L0, E0, b0) in enumerate(batch):
        worst = max(worst, problems.rel_residual(L0, E0, xs[k].cpu().numpy(), b0))
    assert worst < 1e-10, worst
    for k in range(0, 4096, 257):
        L0, E0, b0 = batch[k]
        assert _rel(xs[k].cpu().numpy(), _oracle(L0, E0, b0)[2]) < 1e-11, k
