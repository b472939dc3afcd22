"""Pins of the CPU oracle (oracle/nd.py, oracle/nd_oracle.c) against what the
mathematics fixes, independently of the code under test.

* storage length / level counts: values worked out by hand from the closed
  form sum_{l<L}(N//2^l - 1) (block_tridiag_solver.py:12-23);
* ND schedule and every fill-in index: a brute-force symbolic elimination of
  the block graph;
* factor blocks: the Cholesky factor of the ND-permuted dense matrix is
  unique, so numpy.linalg.cholesky pins every tril(L) and E block;
* solutions: scipy cho_solve, LAPACK banded Cholesky, and the closed form of
  the README generator (Psi = M M^T with M known);
* generators: the reference notebook's printed residual for its own data
  (examples/cholesky_factor.ipynb cell 5: 1.094314663144768e-14);
* the C restatement against the numpy one;
* planted mistakes (in the oracle, on the CPU) that these checks must catch.
"""
import os

import numpy as np
import pytest
import scipy.linalg as la

from oracle import nd, nd_c


# ---------------------------------------------------------------------------
# sizes
# ---------------------------------------------------------------------------

# N -> n_E, by hand: e.g. N=20: 19+9+4+1+0 = 33; N=512: 511+255+...+1+0 = 1013
STORAGE = {1: 0, 2: 1, 3: 2, 4: 4, 5: 5, 7: 8, 8: 11, 20: 33, 100: 190, 128: 247, 200: 389, 512: 1013, 1024: 2036}
LEVELS = {1: 1, 2: 2, 3: 2, 4: 3, 7: 3, 8: 4, 20: 5, 100: 7, 128: 8, 512: 10, 1024: 11}


@pytest.mark.parametrize("N,expect", sorted(STORAGE.items()))
def test_off_diag_storage_len(N, expect):
    assert nd.calculate_off_diag_storage_len(N) == expect


@pytest.mark.parametrize("N,expect", sorted(LEVELS.items()))
def test_recursive_iterations(N, expect):
    assert nd.calculate_recursive_iterations(N) == expect


def test_product_api_sizes_match_oracle():
    from socu_amd.block_tridiag_solver import calculate_off_diag_storage_len, calculate_recursive_iterations

    for N in range(1, 600):
        assert calculate_off_diag_storage_len(N) == nd.calculate_off_diag_storage_len(N)
        assert calculate_recursive_iterations(N) == nd.calculate_recursive_iterations(N)


# ---------------------------------------------------------------------------
# brute-force symbolic elimination of the block graph
# ---------------------------------------------------------------------------

def _symbolic(N):
    """Eliminate the block path graph in ND order.  Returns, per eliminated
    node, its neighbours at elimination time, and the set of fill edges."""
    adj = {i: set() for i in range(N)}
    for i in range(N - 1):
        adj[i].add(i + 1)
        adj[i + 1].add(i)
    neigh = {}
    fills = set()
    for v in nd.nd_order(N):
        nb = sorted(adj[v])
        neigh[v] = nb
        for a in nb:
            for b in nb:
                if a < b and b not in adj[a]:
                    fills.add((a, b))
                    adj[a].add(b)
                    adj[b].add(a)
        for a in nb:
            adj[a].discard(v)
        del adj[v]
    return neigh, fills


@pytest.mark.parametrize("N", list(range(1, 65)) + [100, 129, 200])
def test_schedule_matches_symbolic_elimination(N):
    neigh, fills = _symbolic(N)
    written = {}
    for lev in nd.level_schedule(N):
        s, off, nxt = lev["s"], lev["off"], lev["next_off"]
        for i in lev["nodes"]:
            expect = [j for j in (i - s, i + s) if 0 <= j < N]
            assert neigh[i] == expect, (N, i, neigh[i], expect)
            if i + s < N:
                idx, blk = nd.e_index(N, lev["l"], i // s)
                assert blk == (i + s, i)
                written.setdefault(idx, []).append(("R", i))
            if i >= s:
                idx, blk = nd.e_index(N, lev["l"], (i - s) // s)
                assert blk == (i, i - s)
                written.setdefault(idx, []).append(("Lc", i))
            if i + s < N and i >= s:
                # the fill edge (i-s, i+s) is entry (i-s)//(2s) of level l+1
                f = nxt + (i - s) // (2 * s)
                idx, blk = nd.e_index(N, lev["l"] + 1, (i - s) // (2 * s))
                assert idx == f and blk == (i + s, i - s)
                assert (i - s, i + s) in fills
    # every E index is written by exactly one (node, role) of the factor
    nE = nd.calculate_off_diag_storage_len(N)
    assert sorted(written) == list(range(nE))
    assert all(len(v) == 1 for v in written.values())
    # and the fills are exactly the couplings of the levels above 0
    expected_fills = set()
    for lev in nd.level_schedule(N)[1:]:
        s = lev["s"]
        for k in range(N // s - 1):
            expected_fills.add(((k + 1) * s - 1, (k + 2) * s - 1))
    assert fills == expected_fills


# ---------------------------------------------------------------------------
# factor values and solutions
# ---------------------------------------------------------------------------

CASES = [(1, 1), (1, 2), (2, 3), (3, 4), (4, 5), (5, 7), (3, 8), (8, 9), (6, 20), (4, 33), (2, 64)]


@pytest.mark.parametrize("n,N", CASES)
def test_factor_matches_dense_nd_cholesky(n, N):
    L, E, b, _, _ = nd.readme_problem(n, N, 2, alpha=10.0, seed=N * 7 + n)
    _, _, Lx, Ex = nd.dense_nd_factor(L, E)
    Lf, Ef = L.copy(), E.copy()
    nd.factor(Lf, Ef)
    np.testing.assert_allclose(np.tril(Lf), Lx, rtol=0, atol=1e-11 * max(1, np.abs(Lx).max()))
    np.testing.assert_allclose(Ef, Ex, rtol=0, atol=1e-11 * max(1, np.abs(Ex).max() if Ex.size else 1))


@pytest.mark.parametrize("n,N", CASES)
def test_solve_matches_scipy_and_banded(n, N):
    L, E, b, _, _ = nd.readme_problem(n, N, 3, alpha=10.0, seed=N + 31 * n)
    A = nd.dense_from_blocks(L, E)
    xs = la.cho_solve(la.cho_factor(A, lower=True), b.reshape(N * n, -1))
    # LAPACK banded Cholesky (bandwidth 2n-1) as an independent comparator
    kd = 2 * n - 1
    ab = np.zeros((kd + 1, N * n))
    for j in range(N * n):
        for i in range(j, min(N * n, j + kd + 1)):
            ab[i - j, j] = A[i, j]
    cb = la.cholesky_banded(ab, lower=True)
    xb = la.cho_solve_banded((cb, True), b.reshape(N * n, -1))
    Lf, Ef, x = L.copy(), E.copy(), b.copy()
    nd.factor_and_solve(Lf, Ef, x)
    x = x.reshape(N * n, -1)
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-12
    assert np.linalg.norm(x - xb) / np.linalg.norm(xb) < 1e-12


def test_forward_equals_dense_permuted_triangular_solve():
    n, N = 3, 13
    L, E, b, _, _ = nd.readme_problem(n, N, 2, alpha=10.0, seed=5)
    Lp, order, _, _ = nd.dense_nd_factor(L, E)
    perm = np.concatenate([np.arange(i * n, (i + 1) * n) for i in order])
    y_dense = la.solve_triangular(Lp, b.reshape(N * n, -1)[perm], lower=True)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(N * n)
    Lf, Ef, y = L.copy(), E.copy(), b.copy()
    nd.factor(Lf, Ef)
    nd.forward(Lf, Ef, y)
    np.testing.assert_allclose(y.reshape(N * n, -1), y_dense[inv], atol=1e-12)


def test_readme_generator_closed_form():
    """Psi = M M^T with M = [D_i on the diagonal, Ec_i below]: solve with two
    block-bidiagonal substitutions and compare with the ND solve."""
    n, N, m = 4, 21, 2
    L, E, b, D, Ec = nd.readme_problem(n, N, m, alpha=10.0, seed=0)
    y = np.zeros_like(b)
    for i in range(N):
        rhs = b[i] - (Ec[i - 1] @ y[i - 1] if i > 0 else 0)
        y[i] = la.solve_triangular(D[i], rhs, lower=True)
    x = np.zeros_like(b)
    for i in range(N - 1, -1, -1):
        rhs = y[i] - (Ec[i].T @ x[i + 1] if i < N - 1 else 0)
        x[i] = la.solve_triangular(D[i].T, rhs, lower=False)
    Lf, Ef, xn = L.copy(), E.copy(), b.copy()
    nd.factor_and_solve(Lf, Ef, xn)
    np.testing.assert_allclose(xn, x, rtol=1e-12, atol=1e-14)


def test_notebook_generator_known_answer():
    """The reference notebook prints ||b - A x||_2 = 1.094314663144768e-14 for
    its scipy solution at n=34, N=128 (seed 0).  The same data rebuilt by
    reference_test_problem gives the same residual (up to BLAS rounding)."""
    A, L, E, b = nd.reference_test_problem(34, 128, seed=0)
    Lc = la.cholesky(A, lower=True)
    x = la.solve_triangular(Lc.T, la.solve_triangular(Lc, b.reshape(-1), lower=True), lower=False)
    res = la.norm(b.reshape(-1) - A @ x)
    assert 0.3e-14 < res < 3e-14, res
    # the ND oracle on the same data meets the notebook's reported accuracy
    Lf, Ef, xn = L.copy(), E.copy(), b.copy()
    nd.factor_and_solve(Lf, Ef, xn)
    assert la.norm(xn.reshape(-1) - x) < 1e-14
    assert la.norm(b.reshape(-1) - A @ xn.reshape(-1)) < 1e-13


def test_mask_leaves_inactive_entries_untouched():
    n, N, B = 3, 9, 4
    L0, E0, b0, _, _ = nd.readme_problem(n, N, 1, alpha=10.0, seed=1)
    L = np.tile(L0, (B, 1, 1, 1))
    E = np.tile(E0, (B, 1, 1, 1))
    x = np.tile(b0, (B, 1, 1, 1))
    mask = np.array([True, False, True, False])
    Lc, Ec_, xc = L.copy(), E.copy(), x.copy()
    nd.factor_and_solve(Lc, Ec_, xc, mask)
    for k in range(B):
        if mask[k]:
            assert np.abs(xc[k] - xc[0]).max() == 0
        else:
            assert np.array_equal(Lc[k], L[k]) and np.array_equal(Ec_[k], E[k]) and np.array_equal(xc[k], x[k])


def test_non_spd_gives_nan_silently():
    n, N = 3, 4
    L, E, b, _, _ = nd.readme_problem(n, N, 1, alpha=10.0, seed=0)
    L[2] = -np.eye(n)
    Lf, Ef, x = L.copy(), E.copy(), b.copy()
    nd.factor_and_solve(Lf, Ef, x)
    assert np.isnan(x).any()


# ---------------------------------------------------------------------------
# the C restatement
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("n,N,m", [(1, 1, 1), (3, 2, 2), (8, 7, 3), (5, 20, 1), (32, 33, 2), (17, 100, 4)])
def test_c_oracle_matches_numpy_oracle(n, N, m):
    from oracle import nd_c

    B = 2
    Ls, Es, bs = zip(*[nd.readme_problem(n, N, m, alpha=10.0, seed=s)[:3] for s in range(B)])
    L, E, b = np.stack(Ls), np.stack(Es), np.stack(bs)
    Lc, Ec_, xc = L.copy(), E.copy(), b.copy()
    nd_c.factor_and_solve(Lc, Ec_, xc, np.array([True, True]))
    Ln, En, xn = L.copy(), E.copy(), b.copy()
    nd.factor_and_solve(Ln, En, xn)
    np.testing.assert_allclose(np.tril(Lc), np.tril(Ln), atol=1e-12)
    np.testing.assert_allclose(Ec_, En, atol=1e-12)
    np.testing.assert_allclose(xc, xn, rtol=1e-11, atol=1e-13)


# ---------------------------------------------------------------------------
# planted mistakes: each must be caught by the checks above
# ---------------------------------------------------------------------------

def _factor_with(mistake, L, E):
    """Oracle factor with one deliberate error, for the negative tests."""
    B = L.shape[0]
    N = L.shape[1]
    for lev in nd.level_schedule(N):
        s, off, nxt = lev["s"], lev["off"], lev["next_off"]
        nodes = np.array(lev["nodes"])
        L[:, nodes] = nd._chol(L[:, nodes])
        for i in nodes:
            if i + s < N:
                r = off + i // s
                E[:, r] = nd._T(nd._lower_solve(L[:, i], nd._T(E[:, r])))
            if i >= s:
                c = off + (i - s) // s
                E[:, c] = nd._lower_solve(L[:, i], E[:, c])
        for i in nodes:
            if i + s < N:
                R = E[:, off + i // s]
                L[:, i + s] -= R @ nd._T(R)
            if i >= s:
                C = E[:, off + (i - s) // s]
                L[:, i - s] -= (C @ nd._T(C)) if mistake == "transposed_syrk" else (nd._T(C) @ C)
            if i + s < N and i >= s:
                f = nxt + (i - s) // (2 * s) + (1 if mistake == "fill_index" and i > s else 0)
                f = min(f, E.shape[1] - 1)
                sign = -1.0 if mistake == "fill_sign" else 1.0
                E[:, f] -= sign * (E[:, off + i // s] @ E[:, off + (i - s) // s])


@pytest.mark.parametrize("mistake", ["transposed_syrk", "fill_index", "fill_sign"])
def test_planted_mistakes_are_detected(mistake):
    n, N = 3, 16
    L, E, b, _, _ = nd.readme_problem(n, N, 1, alpha=10.0, seed=2)
    _, _, Lx, Ex = nd.dense_nd_factor(L, E)
    Lf, Ef = L[None].copy(), E[None].copy()
    with np.errstate(all="ignore"):
        _factor_with(mistake, Lf, Ef)
    err = max(np.abs(np.tril(Lf[0]) - Lx).max(), np.abs(Ef[0] - Ex).max())
    assert not (err < 1e-8), f"planted mistake {mistake} went undetected"
    # and the unmodified restatement passes the same check
    Lg, Eg = L[None].copy(), E[None].copy()
    _factor_with("none", Lg, Eg)
    assert max(np.abs(np.tril(Lg[0]) - Lx).max(), np.abs(Eg[0] - Ex).max()) < 1e-10


def test_round_bf16_matches_torch_bfloat16():
    """The BF16 emulation's rounding (FP32, then nearest BF16, ties to even)
    is torch's own float32 -> bfloat16 conversion, bit for bit, including
    ties, powers of two and subnormal-range values."""
    import torch

    rs = np.random.RandomState(0)
    x = np.concatenate([rs.randn(20000) * 10.0 ** rs.randint(-30, 30, 20000),
                        # exact ties: 8 mantissa bits kept, then a half ulp
                        np.float32(1 + np.arange(256) / 128.0 + 1 / 256.0).astype(np.float64),
                        2.0 ** np.arange(-126, 127), [0.0, -0.0, 1e-40, -3e-39]])
    got = nd.round_bf16(x)
    want = torch.tensor(x, dtype=torch.float32).to(torch.bfloat16).double().numpy()
    np.testing.assert_array_equal(got, want)


def test_schur_operand_hook():
    """factor(schur_operand=identity) is the plain factor; BF16 operands move
    the factor by about a BF16 rounding (relative 1e-4 .. 1e-2), not more."""
    L0, E0, _, _, _ = nd.readme_problem(12, 20, 1, alpha=16.0, seed=1)
    La, Ea = L0.copy(), E0.copy()
    nd.factor(La, Ea)
    Lb, Eb = L0.copy(), E0.copy()
    nd.factor(Lb, Eb, schur_operand=lambda a: a)
    np.testing.assert_array_equal(np.tril(La), np.tril(Lb))
    np.testing.assert_array_equal(Ea, Eb)
    Lc, Ec = L0.copy(), E0.copy()
    nd.factor(Lc, Ec, schur_operand=nd.round_bf16)
    d = np.linalg.norm(np.tril(Lc) - np.tril(La)) / np.linalg.norm(np.tril(La))
    assert 1e-5 < d < 1e-2


def test_c_oracle_capi_shape_matches_the_product():
    """The C oracle's ndo_cholesky_* entry points take the argument lists of
    the product's socu_cholesky_* (parsed from include/socu_hip.h and from the
    C source) and compute what its ndo_* functions compute."""
    import re

    from socu_amd import _lib as plib

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

    def params(text, name):
        m = re.search(r"\b" + name + r"\s*\(([^)]*)\)", text)
        assert m, name
        return [re.sub(r"\s+", " ", p.strip()).rsplit(" ", 1)[0].replace(" *", "*") for p in m.group(1).split(",")]

    hdr = re.sub(r"/\*.*?\*/", "", open(os.path.join(root, "include", "socu_hip.h")).read(), flags=re.S)
    src = open(os.path.join(root, "oracle", "nd_oracle.c")).read()
    for name in ("cholesky_factor", "cholesky_forward_substitution", "cholesky_backward_substitution",
                 "cholesky_solve", "cholesky_factor_and_solve"):
        assert params(hdr, "socu_" + name) == params(src, "ndo_" + name), name
    lib = nd_c.capi()
    n, N, m, B = 6, 13, 2, 2
    L0 = np.stack([nd.readme_problem(n, N, m, alpha=10.0, seed=s)[0] for s in range(B)])
    E0 = np.stack([nd.readme_problem(n, N, m, alpha=10.0, seed=s)[1] for s in range(B)])
    b0 = np.stack([nd.readme_problem(n, N, m, alpha=10.0, seed=s)[2] for s in range(B)])
    L, E, x = L0.copy(), E0.copy(), b0.copy()
    assert lib.ndo_cholesky_factor_and_solve(L.ctypes.data, E.ctypes.data, x.ctypes.data, None, B, N, n, m,
                                             plib.DTYPE_F64, None) == 0
    Lr, Er, xr = L0.copy(), E0.copy(), b0.copy()
    nd_c.factor_and_solve(Lr, Er, xr)
    np.testing.assert_array_equal(x, xr)
    y = b0.copy()
    assert lib.ndo_cholesky_forward_substitution(L.ctypes.data, E.ctypes.data, y.ctypes.data, None, B, N, n, m,
                                                 plib.DTYPE_F64, None, None) == 0
    assert lib.ndo_cholesky_backward_substitution(L.ctypes.data, E.ctypes.data, y.ctypes.data, None, B, N, n, m,
                                                  plib.DTYPE_F64, None, None) == 0
    np.testing.assert_allclose(y, xr, rtol=1e-13, atol=0)
    assert lib.ndo_cholesky_factor(L.ctypes.data, E.ctypes.data, None, B, N, n, plib.DTYPE_F32, None) == plib.EUNSUPPORTED
    assert lib.ndo_cholesky_factor(None, None, None, 1, N, n, plib.DTYPE_F64, None) == plib.EINVAL
