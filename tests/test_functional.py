"""The functional front-end (socu_amd.functional), the counterpart of socu's
JAX interface (socu/jax.py:270-337), on the GPU.

* The reference's JAX test grid (tests/test_jax.py:29-146): N x pad_problem
  x n x dtype, E passed as the N-1 coupling blocks (short of the storage
  length), separate factor + solve, factor_and_solve, forward + backward,
  with the reference's absolute tolerances against scipy.
* Its two vmap tests (tests/test_jax.py:148-213), through torch.func.vmap.
* Against the oracle: factors (tril(L) and every E index) and solutions.
* Inputs are never modified; a padded n and leading batch dimensions.
* Gradients: factor_and_solve's x against autograd through a dense FP64
  torch.linalg.solve on the CPU; solve / forward / backward against autograd
  through dense triangular solves with the ND-permuted factor assembled from
  the storage on the CPU, and against central finite differences of the
  oracle's own substitutions.
"""
import numpy as np
import pytest
import scipy.linalg as la

from oracle import nd

pytestmark = pytest.mark.gpu

TOL = {"float32": 1e-4, "float64": 1e-10}


def _ref_data(n, N):
    """tests/test_jax.py:29-68: blocks of A = Lf Lf^T, E as N-1 blocks,
    b = rand(N*n), x_ref by scipy triangular solves."""
    A, L, E, b = nd.reference_test_problem(n, N, seed=42)
    Lc = la.cholesky(A, lower=True)
    y = la.solve_triangular(Lc, b.reshape(-1), lower=True)
    x_ref = la.solve_triangular(Lc.T, y, lower=False)
    return A, L, E[: N - 1], b, x_ref


@pytest.mark.parametrize("dtype_name", ["float32", "float64"])
@pytest.mark.parametrize("n", [1, 10, 32, 36, 64, 68])
@pytest.mark.parametrize("pad_problem", [True, False])
@pytest.mark.parametrize("N", [1, 2, 3, 4, 7, 8, 20, 100])
def test_reference_jax_grid(N, pad_problem, n, dtype_name, device):
    import torch

    from socu_amd.functional import (
        backward_substitution,
        cholesky_factor,
        cholesky_factor_and_solve,
        cholesky_solve,
        forward_substitution,
    )

    tol = TOL[dtype_name]
    dtype = getattr(torch, dtype_name)
    A, L0, E0, b0, x_ref = _ref_data(n, N)
    L = torch.tensor(L0, dtype=dtype, device=device)
    E = torch.tensor(E0, dtype=dtype, device=device)
    b = torch.tensor(b0, dtype=dtype, device=device)
    Lf, Ef = cholesky_factor(L, E, pad_problem=pad_problem)
    assert Ef.shape[0] == nd.calculate_off_diag_storage_len(N)
    x = cholesky_solve(Lf, Ef, b, pad_problem=pad_problem).double().cpu().numpy().ravel()
    assert la.norm(x_ref - x) < tol
    assert la.norm(b0.ravel() - A @ x) < tol
    _, _, xc = cholesky_factor_and_solve(L, E, b, pad_problem=pad_problem)
    xc = xc.double().cpu().numpy().ravel()
    assert la.norm(x_ref - xc) < tol
    assert la.norm(b0.ravel() - A @ xc) < tol
    assert la.norm(x - xc) < tol
    y = forward_substitution(Lf, Ef, b, pad_problem=pad_problem)
    xs = backward_substitution(Lf, Ef, y, pad_problem=pad_problem).double().cpu().numpy().ravel()
    assert la.norm(x - xs) < tol
    # the inputs are untouched
    np.testing.assert_array_equal(L.cpu().numpy(), L0.astype(dtype_name))
    np.testing.assert_array_equal(E.cpu().numpy(), E0.astype(dtype_name))
    np.testing.assert_array_equal(b.cpu().numpy(), b0.astype(dtype_name))


@pytest.mark.parametrize("pad_problem", [True, False])
def test_reference_vmap_factor_and_solve(pad_problem, device):
    """tests/test_jax.py:148-185 through torch.func.vmap."""
    import torch

    from socu_amd.functional import cholesky_factor_and_solve

    n, N, batch = 4, 3, 2
    Ls, Es, bs, xr = [], [], [], []
    for seed in range(batch):
        A, L0, E0, b0, _ = _ref_data(n, N)
        Ls.append(L0 + 0.1 * seed * np.eye(n)[None])
        Es.append(E0)
        bs.append(b0)
        As = A.copy()
        for i in range(N):
            As[i * n:(i + 1) * n, i * n:(i + 1) * n] += 0.1 * seed * np.eye(n)
        xr.append(la.solve(As, b0.reshape(N * n, 1)).reshape(N, n, 1))
    Lb = torch.tensor(np.stack(Ls), device=device)
    Eb = torch.tensor(np.stack(Es), device=device)
    bb = torch.tensor(np.stack(bs), device=device)
    _, _, x = torch.func.vmap(lambda L, E, b: cholesky_factor_and_solve(L, E, b, pad_problem=pad_problem))(Lb, Eb, bb)
    assert la.norm(x.cpu().numpy() - np.stack(xr)) < 1e-10


@pytest.mark.parametrize("pad_problem", [True, False])
def test_reference_vmap_rhs_only(pad_problem, device):
    """tests/test_jax.py:187-213: vmap over right-hand sides, shared factor."""
    import torch

    from socu_amd.functional import cholesky_factor, cholesky_solve

    n, N = 4, 3
    A, L0, E0, b0, _ = _ref_data(n, N)
    Lf, Ef = cholesky_factor(torch.tensor(L0, device=device), torch.tensor(E0, device=device), pad_problem=pad_problem)
    b = torch.tensor(b0, device=device)
    bb = torch.stack([b, 2.0 * b])
    x = torch.func.vmap(lambda bv: cholesky_solve(Lf, Ef, bv, pad_problem=pad_problem))(bb)
    xr = la.solve(A, b0.reshape(N * n, 1)).reshape(N, n, 1)
    assert la.norm(x.cpu().numpy() - np.stack([xr, 2.0 * xr])) < 1e-10


@pytest.mark.parametrize("n,N,m,lead", [(8, 13, 2, (3,)), (36, 20, 1, (2, 2)), (64, 9, 3, ())])
def test_factor_and_substitutions_vs_oracle(n, N, m, lead, device):
    import torch

    from socu_amd.functional import backward_substitution, cholesky_factor, forward_substitution

    count = int(np.prod(lead)) if lead else 1
    L0, E0, b0 = [], [], []
    for k in range(count):
        l_, e_, b_, _, _ = nd.readme_problem(n, N, m, alpha=16.0, seed=70 + k)
        L0.append(l_)
        E0.append(e_)
        b0.append(b_)
    L0, E0, b0 = np.stack(L0), np.stack(E0), np.stack(b0)
    Lr, Er, yr = L0.copy(), E0.copy(), b0.copy()
    nd.factor(Lr, Er)
    nd.forward(Lr, Er, yr)
    xr = yr.copy()
    nd.backward(Lr, Er, xr)
    shp = lambda a: a.reshape(*lead, *a.shape[1:]) if lead else a[0]  # noqa: E731
    L = torch.tensor(shp(L0), device=device)
    E = torch.tensor(shp(E0[:, : N - 1]), device=device)  # short E
    b = torch.tensor(shp(b0), device=device)
    Lf, Ef = cholesky_factor(L, E)
    assert Lf.shape == L.shape and Ef.shape[-3] == nd.calculate_off_diag_storage_len(N)
    Lg = Lf.reshape(-1, N, n, n).cpu().numpy()
    Eg = Ef.reshape(-1, Ef.shape[-3], n, n).cpu().numpy()
    rel = lambda a, r: np.linalg.norm(a - r) / np.linalg.norm(r)  # noqa: E731
    assert rel(np.tril(Lg), np.tril(Lr)) < 1e-11
    assert rel(Eg, Er) < 1e-11
    y = forward_substitution(Lf, Ef, b)
    assert rel(y.reshape(yr.shape).cpu().numpy(), yr) < 1e-11
    x = backward_substitution(Lf, Ef, y)
    assert rel(x.reshape(xr.shape).cpu().numpy(), xr) < 1e-11


def test_padded_block_size(device):
    """n = 36 in FP32 pads to 48, n = 10 in FP64 to 16: the real blocks are
    unchanged, the results equal the unpadded ones to rounding."""
    import torch

    from socu_amd.functional import cholesky_factor_and_solve

    for n, dtype, tol in ((36, torch.float32, 2e-4), (10, torch.float64, 1e-12)):
        L0, E0, b0, _, _ = nd.readme_problem(n, 17, 2, alpha=16.0, seed=3)
        L = torch.tensor(L0, dtype=dtype, device=device)
        E = torch.tensor(E0, dtype=dtype, device=device)
        b = torch.tensor(b0, dtype=dtype, device=device)
        Lp, Ep, xp = cholesky_factor_and_solve(L, E, b, pad_problem=True)
        Lu, Eu, xu = cholesky_factor_and_solve(L, E, b, pad_problem=False)
        assert Lp.shape == L.shape and xp.shape == b.shape
        rel = lambda a, r: (torch.linalg.norm(a - r) / torch.linalg.norm(r)).item()  # noqa: E731
        assert rel(xp, xu) < tol
        assert rel(torch.tril(Lp), torch.tril(Lu)) < tol
        assert rel(Ep, Eu) < tol


def test_validation_messages(device):
    import torch

    from socu_amd.functional import cholesky_factor, cholesky_solve

    L = torch.zeros(4, 3, 3, dtype=torch.float64, device=device)
    E = torch.zeros(3, 3, 3, dtype=torch.float64, device=device)
    with pytest.raises(ValueError, match="3- or 4-dimensional"):
        cholesky_factor(L[0], E[0])
    with pytest.raises(ValueError, match="same rank"):
        cholesky_factor(L, E[None])
    with pytest.raises(ValueError, match="square"):
        cholesky_factor(torch.zeros(4, 3, 2, dtype=torch.float64, device=device), E)
    with pytest.raises(ValueError, match="horizon dimension"):
        cholesky_solve(L, E, torch.zeros(5, 3, 1, dtype=torch.float64, device=device))
    with pytest.raises(ValueError, match="same dtype"):
        cholesky_solve(L, E, torch.zeros(4, 3, 1, dtype=torch.float32, device=device))
    with pytest.raises(ValueError, match="no CPU path"):
        cholesky_factor(L.cpu(), E.cpu())


# ---------------------------------------------------------------------------
# gradients
# ---------------------------------------------------------------------------

def _dense_psi_torch(L, E):
    import torch

    N, n, _ = L.shape
    A = torch.zeros(N * n, N * n, dtype=L.dtype)
    for i in range(N):
        Li = torch.tril(L[i])
        A[i * n:(i + 1) * n, i * n:(i + 1) * n] = Li + torch.tril(Li, -1).T
        if i < N - 1:
            A[(i + 1) * n:(i + 2) * n, i * n:(i + 1) * n] = E[i]
            A[i * n:(i + 1) * n, (i + 1) * n:(i + 2) * n] = E[i].T
    return A


@pytest.mark.parametrize("n,N,m,pad", [(5, 9, 2, True), (32, 12, 3, False)])
def test_grad_factor_and_solve_vs_dense(n, N, m, pad, device):
    import torch

    from socu_amd.functional import cholesky_factor_and_solve

    L0, E0, b0, _, _ = nd.readme_problem(n, N, m, alpha=16.0, seed=8)
    w = np.random.RandomState(1).randn(N, n, m)
    L = torch.tensor(L0, device=device, requires_grad=True)
    E = torch.tensor(E0[: N - 1], device=device, requires_grad=True)
    b = torch.tensor(b0, device=device, requires_grad=True)
    _, _, x = cholesky_factor_and_solve(L, E, b, pad_problem=pad)
    (x * torch.tensor(w, device=device)).sum().backward()
    Lc = torch.tensor(L0, requires_grad=True)
    Ec = torch.tensor(E0[: N - 1], requires_grad=True)
    bc = torch.tensor(b0, requires_grad=True)
    xc = torch.linalg.solve(_dense_psi_torch(Lc, Ec), bc.reshape(N * n, m)).reshape(N, n, m)
    (xc * torch.tensor(w)).sum().backward()
    for g, r in ((L.grad, Lc.grad), (E.grad, Ec.grad), (b.grad, bc.grad)):
        g = g.cpu().numpy()
        r = r.numpy()
        assert np.linalg.norm(g - r) / np.linalg.norm(r) < 1e-10


def _dense_factor_solves(Lf, Ef, b, mode):
    """forward / backward / solve with the ND factor assembled densely in the
    elimination order from the storage (torch, differentiable in Lf, Ef)."""
    import torch

    N, n, _ = Lf.shape
    order = nd.nd_order(N)
    pos = {node: p for p, node in enumerate(order)}
    Lp = torch.zeros(N * n, N * n, dtype=Lf.dtype)
    for i in range(N):
        p = pos[i]
        Lp[p * n:(p + 1) * n, p * n:(p + 1) * n] = torch.tril(Lf[i])
    for lev in nd.level_schedule(N):
        s, off = lev["s"], lev["off"]
        for k in range(N // s - 1):
            a, bb = (k + 1) * s - 1, (k + 2) * s - 1
            pa, pb = pos[a], pos[bb]
            if pa < pb:
                Lp[pb * n:(pb + 1) * n, pa * n:(pa + 1) * n] = Ef[off + k]
            else:
                Lp[pa * n:(pa + 1) * n, pb * n:(pb + 1) * n] = Ef[off + k].T
    perm = torch.tensor(np.concatenate([np.arange(i * n, (i + 1) * n) for i in order]))
    m = b.shape[-1]
    v = b.reshape(N * n, m)
    if mode in ("forward", "solve"):
        yp = torch.linalg.solve_triangular(Lp, v[perm], upper=False)
        v = torch.zeros_like(v).index_copy(0, perm, yp)
    if mode in ("backward", "solve"):
        xp = torch.linalg.solve_triangular(Lp.T.contiguous(), v[perm], upper=True)
        v = torch.zeros_like(v).index_copy(0, perm, xp)
    return v.reshape(N, n, m)


@pytest.mark.parametrize("mode", ["forward", "backward", "solve"])
@pytest.mark.parametrize("n,N,m", [(4, 7, 2), (33, 10, 1)])
def test_grad_substitutions_vs_dense_factor(mode, n, N, m, device):
    import torch

    from socu_amd import functional as fn

    L0, E0, b0, _, _ = nd.readme_problem(n, N, m, alpha=16.0, seed=21)
    Lf0, Ef0 = L0.copy(), E0.copy()
    nd.factor(Lf0, Ef0)
    Lf0 = np.tril(Lf0)
    w = np.random.RandomState(2).randn(N, n, m)
    f = {"forward": fn.forward_substitution, "backward": fn.backward_substitution, "solve": fn.cholesky_solve}[mode]
    Lf = torch.tensor(Lf0, device=device, requires_grad=True)
    Ef = torch.tensor(Ef0, device=device, requires_grad=True)
    b = torch.tensor(b0, device=device, requires_grad=True)
    out = f(Lf, Ef, b, pad_problem=False)
    (out * torch.tensor(w, device=device)).sum().backward()
    Lc = torch.tensor(Lf0, requires_grad=True)
    Ec = torch.tensor(Ef0, requires_grad=True)
    bc = torch.tensor(b0, requires_grad=True)
    oc = _dense_factor_solves(Lc, Ec, bc, mode)
    np.testing.assert_allclose(out.detach().cpu().numpy(), oc.detach().numpy(), rtol=0, atol=1e-11)
    (oc * torch.tensor(w)).sum().backward()
    for g, r in ((Lf.grad, Lc.grad), (Ef.grad, Ec.grad), (b.grad, bc.grad)):
        g = g.cpu().numpy()
        r = r.numpy()
        assert np.linalg.norm(g - r) / np.linalg.norm(r) < 1e-10


def test_grad_solve_vs_oracle_finite_differences(device):
    """Central differences (h = 1e-6) of the oracle's forward + backward
    substitution with respect to 40 random factor entries."""
    import torch

    from socu_amd.functional import cholesky_solve

    n, N, m = 3, 6, 2
    L0, E0, b0, _, _ = nd.readme_problem(n, N, m, alpha=16.0, seed=4)
    Lf0, Ef0 = L0.copy(), E0.copy()
    nd.factor(Lf0, Ef0)
    w = np.random.RandomState(3).randn(N, n, m)
    Lf = torch.tensor(Lf0, device=device, requires_grad=True)
    Ef = torch.tensor(Ef0, device=device, requires_grad=True)
    x = cholesky_solve(Lf, Ef, torch.tensor(b0, device=device), pad_problem=False)
    (x * torch.tensor(w, device=device)).sum().backward()
    gL, gE = Lf.grad.cpu().numpy(), Ef.grad.cpu().numpy()

    def loss(Lx, Ex):
        xx = b0.copy()
        nd.forward(Lx, Ex, xx)
        nd.backward(Lx, Ex, xx)
        return float((xx * w).sum())

    rs = np.random.RandomState(9)
    h = 1e-6
    for _ in range(40):
        if rs.rand() < 0.5:
            i, r = rs.randint(N), rs.randint(n)
            c = rs.randint(r + 1)
            Lp_, Lm_ = Lf0.copy(), Lf0.copy()
            Lp_[i, r, c] += h
            Lm_[i, r, c] -= h
            fd = (loss(Lp_, Ef0) - loss(Lm_, Ef0)) / (2 * h)
            g = gL[i, r, c]
        else:
            e, r, c = rs.randint(Ef0.shape[0]), rs.randint(n), rs.randint(n)
            Ep_, Em_ = Ef0.copy(), Ef0.copy()
            Ep_[e, r, c] += h
            Em_[e, r, c] -= h
            fd = (loss(Lf0, Ep_) - loss(Lf0, Em_)) / (2 * h)
            g = gE[e, r, c]
        assert abs(g - fd) <= 1e-7 * max(1.0, abs(fd)), (g, fd)
    # the strict upper triangle of the L blocks is not read: zero gradient
    assert np.all(np.triu(gL, 1) == 0)
