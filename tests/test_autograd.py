"""Gradients of the differentiable solve x = Psi(L, E)^{-1} b.

Checked three independent ways: against autograd through a dense FP64
``torch.linalg.solve`` on the CPU (the matrix assembled from the same blocks),
against central finite differences of the GPU forward along random
directions, and with ``torch.autograd.gradcheck`` (element-wise finite
differences) on a tiny system.  Tolerances are stated per check.
"""
import numpy as np
import pytest

from oracle import nd

pytestmark = pytest.mark.gpu


def _dense_psi(L, E):
    """Dense Psi from blocks, differentiable, in torch (CPU reference)."""
    import torch

    N, n = L.shape[0], L.shape[1]
    Ls = torch.tril(L) + torch.tril(L, -1).transpose(-1, -2)
    zero = torch.zeros(n, n, dtype=L.dtype)
    rows = []
    for i in range(N):
        row = []
        for j in range(N):
            if i == j:
                row.append(Ls[i])
            elif i == j + 1:
                row.append(E[j])
            elif j == i + 1:
                row.append(E[i].transpose(0, 1))
            else:
                row.append(zero)
        rows.append(torch.cat(row, 1))
    return torch.cat(rows, 0)


def _dense_grads(L0, E0, b0, w):
    import torch

    N, n, m = b0.shape
    L = torch.tensor(L0, requires_grad=True)
    E = torch.tensor(E0, requires_grad=True)
    b = torch.tensor(b0, requires_grad=True)
    x = torch.linalg.solve(_dense_psi(L, E), b.reshape(N * n, m)).reshape(N, n, m)
    (x * torch.tensor(w)).sum().backward()
    gE = E.grad.numpy() if E.grad is not None else np.zeros_like(E0)  # N = 1: E is empty
    return x.detach().numpy(), L.grad.numpy(), gE, b.grad.numpy()


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("n,N,m", [(2, 1, 1), (3, 5, 2), (8, 9, 1), (16, 20, 4), (32, 33, 3), (5, 16, 2)])
def test_matches_dense_torch_gradient(n, N, m, device):
    import torch
    from socu_amd.autograd import block_tridiag_solve

    L0, E0, b0, _, _ = nd.readme_problem(n, N, m, alpha=10.0, seed=n + N)
    # non-zero upper triangle and fill-in entries: both must be ignored
    L0 = L0 + np.triu(np.random.RandomState(1).randn(N, n, n), 1)
    if N > 1:
        E0 = E0.copy()
        E0[N - 1:] = np.random.RandomState(6).randn(*E0[N - 1:].shape)
    w = np.random.RandomState(2).randn(N, n, m)
    xd, gLd, gEd, gbd = _dense_grads(L0, E0, b0, w)
    L = torch.tensor(L0, device=device, requires_grad=True)
    E = torch.tensor(E0, device=device, requires_grad=True)
    b = torch.tensor(b0, device=device, requires_grad=True)
    x = block_tridiag_solve(L, E, b)
    (x * torch.tensor(w, device=device)).sum().backward()
    assert _rel(x.detach().cpu().numpy(), xd) < 1e-11
    assert _rel(b.grad.cpu().numpy(), gbd) < 1e-11
    # the dense path differentiates Psi through tril(L) mirrored: same parameterisation
    assert _rel(L.grad.cpu().numpy(), gLd) < 1e-11
    assert _rel(E.grad.cpu().numpy(), gEd) < 1e-11
    assert np.all(np.triu(L.grad.cpu().numpy(), 1) == 0)
    assert np.all(E.grad.cpu().numpy()[max(N - 1, 0):] == 0)
    np.testing.assert_array_equal(L.detach().cpu().numpy(), L0)  # inputs untouched


def test_ignores_upper_triangle_and_fill_in(device):
    import torch
    from socu_amd.autograd import block_tridiag_solve

    L0, E0, b0, _, _ = nd.readme_problem(8, 12, 2, alpha=10.0, seed=4)
    rs = np.random.RandomState(3)
    L1 = L0 + np.triu(rs.randn(*L0.shape), 1)
    E1 = E0.copy()
    E1[11:] = rs.randn(*E1[11:].shape)
    x0 = block_tridiag_solve(*(torch.tensor(a, device=device) for a in (L0, E0, b0)))
    x1 = block_tridiag_solve(*(torch.tensor(a, device=device) for a in (L1, E1, b0)))
    assert torch.equal(x0, x1)


def test_central_finite_differences(device):
    import torch
    from socu_amd.autograd import block_tridiag_solve

    n, N, m = 12, 17, 3
    L0, E0, b0, _, _ = nd.readme_problem(n, N, m, alpha=10.0, seed=7)
    rs = np.random.RandomState(8)
    w = torch.tensor(rs.randn(N, n, m), device=device)
    L = torch.tensor(L0, device=device, requires_grad=True)
    E = torch.tensor(E0, device=device, requires_grad=True)
    b = torch.tensor(b0, device=device, requires_grad=True)
    (block_tridiag_solve(L, E, b) * w).sum().backward()

    def f(dL, dE, db, t):
        with torch.no_grad():
            return float((block_tridiag_solve(L + t * dL, E + t * dE, b + t * db) * w).sum())

    for trial in range(4):
        dL = torch.tensor(np.tril(rs.randn(N, n, n)), device=device)
        dE = torch.zeros_like(E)
        dE[: N - 1] = torch.tensor(rs.randn(N - 1, n, n), device=device)
        db = torch.tensor(rs.randn(N, n, m), device=device)
        h = 1e-5
        fd = (f(dL, dE, db, h) - f(dL, dE, db, -h)) / (2 * h)
        an = float((L.grad * dL).sum() + (E.grad * dE).sum() + (b.grad * db).sum())
        assert abs(fd - an) <= 1e-7 * max(1.0, abs(an)), (trial, fd, an)


def test_gradcheck_tiny(device):
    import torch
    from socu_amd.autograd import block_tridiag_solve

    L0, E0, b0, _, _ = nd.readme_problem(2, 4, 2, alpha=3.0, seed=9)
    L = torch.tensor(L0, device=device, requires_grad=True)
    E = torch.tensor(E0, device=device, requires_grad=True)
    b = torch.tensor(b0, device=device, requires_grad=True)
    assert torch.autograd.gradcheck(block_tridiag_solve, (L, E, b), eps=1e-6, atol=1e-8, rtol=1e-6)


def test_batched_4d_and_float32(device):
    import torch
    from socu_amd.autograd import block_tridiag_solve

    n, N, m, B = 8, 10, 2, 3
    probs = [nd.readme_problem(n, N, m, alpha=10.0, seed=20 + k) for k in range(B)]
    L0 = np.stack([p[0] for p in probs])
    E0 = np.stack([p[1] for p in probs])
    b0 = np.stack([p[2] for p in probs])
    w = np.random.RandomState(5).randn(B, N, n, m)
    for dtype, tol in ((torch.float64, 1e-11), (torch.float32, 2e-4)):
        L = torch.tensor(L0, device=device, dtype=dtype, requires_grad=True)
        E = torch.tensor(E0, device=device, dtype=dtype, requires_grad=True)
        b = torch.tensor(b0, device=device, dtype=dtype, requires_grad=True)
        (block_tridiag_solve(L, E, b) * torch.tensor(w, device=device, dtype=dtype)).sum().backward()
        for k in range(B):
            _, gL, gE, gb = _dense_grads(L0[k], E0[k], b0[k], w[k])
            assert _rel(L.grad[k].double().cpu().numpy(), gL) < tol, (dtype, k)
            assert _rel(E.grad[k].double().cpu().numpy(), gE) < tol, (dtype, k)
            assert _rel(b.grad[k].double().cpu().numpy(), gb) < tol, (dtype, k)


def test_c5_shape_sampled_dense(device):
    """B=256, n=32, N=128, m=4 (the gradient workload); two systems checked
    against the dense FP64 gradient (4096 x 4096 solves on the CPU)."""
    import torch
    from socu_amd.autograd import block_tridiag_solve

    n, N, m, B = 32, 128, 4, 256
    gen = torch.Generator(device="cpu").manual_seed(2)
    Dd = torch.tril(torch.randn(B, N, n, n, generator=gen, dtype=torch.float64)) + 10 * torch.eye(n, dtype=torch.float64)
    Ec = torch.randn(B, N - 1, n, n, generator=gen, dtype=torch.float64)
    L0 = Dd @ Dd.transpose(-1, -2)
    L0[:, 1:] += Ec @ Ec.transpose(-1, -2)
    nE = nd.calculate_off_diag_storage_len(N)
    E0 = torch.zeros(B, nE, n, n, dtype=torch.float64)
    E0[:, : N - 1] = Ec @ Dd[:, : N - 1].transpose(-1, -2)
    b0 = torch.randn(B, N, n, m, generator=gen, dtype=torch.float64)
    w = torch.randn(B, N, n, m, generator=gen, dtype=torch.float64)
    L = L0.to(device).requires_grad_(True)
    E = E0.to(device).requires_grad_(True)
    b = b0.to(device).requires_grad_(True)
    (block_tridiag_solve(L, E, b) * w.to(device)).sum().backward()
    for k in (0, 255):
        _, gL, gE, gb = _dense_grads(L0[k].numpy(), E0[k].numpy(), b0[k].numpy(), w[k].numpy())
        assert _rel(L.grad[k].cpu().numpy(), gL) < 1e-10, k
        assert _rel(E.grad[k].cpu().numpy(), gE) < 1e-10, k
        assert _rel(b.grad[k].cpu().numpy(), gb) < 1e-10, k
