"""CPU oracle for the nested-dissection (ND) block-tridiagonal Cholesky path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s CPU-baseline leg may import this module, and only as the checker.
The product (``socu_amd``) never imports it and has no CPU fallback.

This module restates, in float64 numpy, the algorithm of socu's
``socu/block_tridiag_solver.py`` (PREDICT-EPFL/socu) from the maths, with the
same elimination order, the same E storage layout (including every fill-in
index) and the same per-node operations.  Each function cites the reference
lines it follows.

How each function is pinned (checked by ``tests/test_oracle.py``):

* ``calculate_recursive_iterations`` / ``calculate_off_diag_storage_len``:
  the table of values worked out by hand from the closed form
  sum_{l<L}(N//2^l - 1), and a brute-force count of the couplings per level.
* ``level_schedule`` / ``e_index``: a brute-force symbolic elimination of the
  block graph (every fill-in lands exactly at the E index the schedule names,
  every E index is written once).
* ``factor``: the Cholesky factor is unique for SPD input, so every output
  block is compared with ``numpy.linalg.cholesky`` of the ND-permuted dense
  matrix (``dense_nd_factor``), and with LAPACK banded Cholesky through the
  solution.
* ``forward`` / ``backward`` / ``solve``: the dense permuted triangular solves,
  ``scipy.linalg.cho_solve``, and the closed-form block-bidiagonal solve of the
  README generator (Psi = M M^T with M known).
* The generators: the README generator's own residual check
  (README.md:146-156) and the reference test's data builder.
"""
from __future__ import annotations

import numpy as np

__all__ = [
    "calculate_recursive_iterations",
    "calculate_off_diag_storage_len",
    "level_schedule",
    "e_index",
    "nd_order",
    "readme_batch",
    "readme_problem",
    "reference_test_problem",
    "dense_from_blocks",
    "factor",
    "forward",
    "backward",
    "solve",
    "factor_and_solve",
    "dense_nd_factor",
    "block_tridiag_matvec",
    "rel_residual",
    "nd_flops",
    "nd_bytes",
]


# ---------------------------------------------------------------------------
# sizes and schedule  (block_tridiag_solver.py:12-23, :793-813, :869-914)
# ---------------------------------------------------------------------------

def calculate_recursive_iterations(N: int) -> int:
    """Number of ND levels, floor(log2 N) + 1 (block_tridiag_solver.py:12-15)."""
    assert N > 0
    return int(N).bit_length()


def calculate_off_diag_storage_len(N: int) -> int:
    """Blocks in the E storage: sum over levels of N//2^l - 1 (:18-23)."""
    return sum(N // (1 << l) - 1 for l in range(calculate_recursive_iterations(N)))


def level_schedule(N: int):
    """Per-level parameters of the ND elimination.

    Returns a list of dicts with keys ``l, s, dim, off, next_off, nodes``:
    nodes of level l are i = s-1+2*s*t for t < dim = (N+s)//(2s)
    (block_tridiag_solver.py:174, :798, :873); ``off`` is the first E index of
    the level's segment and ``next_off`` the first of the next level's segment
    (:869-871, :912-914).
    """
    out = []
    off = 0
    for l in range(calculate_recursive_iterations(N)):
        s = 1 << l
        dim = (N + s) // (2 * s)
        nxt = off + N // s - 1
        out.append(dict(l=l, s=s, dim=dim, off=off, next_off=nxt,
                        nodes=[s - 1 + 2 * s * t for t in range(dim)]))
        off = nxt
    return out


def e_index(N: int, l: int, k: int):
    """E storage index of coupling k of level l, and the block it holds.

    Level-l segment entry k couples column a=(k+1)s-1 and row b=(k+2)s-1
    (block (b, a) of Psi before the factor).  Index maps used by the kernels:
    R of node i is k=i//s (:114, :239), Lc of node i is k=(i-s)//s (:133, :270),
    the fill-in F of node i is entry (i-s)//(2s) of level l+1 (:145, :367).
    """
    sched = level_schedule(N)
    s = 1 << l
    assert 0 <= k <= N // s - 2
    return sched[l]["off"] + k, ((k + 2) * s - 1, (k + 1) * s - 1)


def nd_order(N: int):
    """Elimination order P: concatenation of the level node lists."""
    order = []
    for lev in level_schedule(N):
        order.extend(lev["nodes"])
    assert sorted(order) == list(range(N))
    return order


# ---------------------------------------------------------------------------
# problem generators
# ---------------------------------------------------------------------------

def readme_problem(n: int, N: int, m: int = 1, alpha: float = 10.0, seed: int = 0):
    """The README generator (README.md:100-121), seeded.

    D_i = tril(randn(n,n)) + alpha*I, Ec_i = randn(n,n) drawn interleaved per i,
    then b = randn(N,n,m).  L[i] = D_i D_i^T + Ec_{i-1} Ec_{i-1}^T,
    E[i] = Ec_i D_i^T, E zero-padded to calculate_off_diag_storage_len(N).
    Psi = M M^T where M is lower block bidiagonal with D_i on the diagonal and
    Ec_i below it, which the closed-form check uses.

    Returns (L, E, b, D, Ec).
    """
    rs = np.random.RandomState(seed)
    D = np.zeros((N, n, n))
    Ec = np.zeros((max(N - 1, 0), n, n))
    for i in range(N):
        D[i] = np.tril(rs.randn(n, n)) + alpha * np.eye(n)
        if i < N - 1:
            Ec[i] = rs.randn(n, n)
    b = rs.randn(N, n, m)
    L = np.einsum("nij,nkj->nik", D, D)
    if N > 1:
        L[1:] += np.einsum("nij,nkj->nik", Ec, Ec)
    E = np.zeros((calculate_off_diag_storage_len(N), n, n))
    if N > 1:
        E[: N - 1] = np.einsum("nij,nkj->nik", Ec, D[: N - 1])
    return L, E, b, D, Ec


def readme_batch(count: int, n: int, N: int, m: int = 1, alpha: float = 10.0, seed: int = 0,
                 with_weights: bool = False):
    """``count`` systems of the README generator drawn from one stream
    RandomState(seed), each system's draws in turn (the generator's own order:
    D_i, Ec_i interleaved, then b), then, if ``with_weights``, a loss weight w
    of b's shape per system.  Returns stacked (L, E, b[, w]) with a leading
    batch axis."""
    rs = np.random.RandomState(seed)
    Ls, Es, bs = [], [], []
    for _ in range(count):
        D = np.zeros((N, n, n))
        Ec = np.zeros((max(N - 1, 0), n, n))
        for i in range(N):
            D[i] = np.tril(rs.randn(n, n)) + alpha * np.eye(n)
            if i < N - 1:
                Ec[i] = rs.randn(n, n)
        b = rs.randn(N, n, m)
        L = np.einsum("nij,nkj->nik", D, D)
        if N > 1:
            L[1:] += np.einsum("nij,nkj->nik", Ec, Ec)
        E = np.zeros((calculate_off_diag_storage_len(N), n, n))
        if N > 1:
            E[: N - 1] = np.einsum("nij,nkj->nik", Ec, D[: N - 1])
        Ls.append(L)
        Es.append(E)
        bs.append(b)
    out = [np.stack(Ls), np.stack(Es), np.stack(bs)]
    if with_weights:
        out.append(rs.randn(count, N, n, m))
    return tuple(out)


def reference_test_problem(n: int, N: int, seed: int = 42):
    """The data builder of the reference's tests
    (tests/test_block_tridiag_solver.py:17-52): a random block-bidiagonal factor
    tril(rand)+10I / rand, A = Lf Lf^T, blocks extracted from A, b = rand(N*n)
    drawn after re-seeding.  Returns (A, L, E, b) with b shaped (N, n, 1)."""
    np.random.seed(seed)
    Lf = np.zeros((N * n, N * n))
    for i in range(N):
        Lf[i * n:(i + 1) * n, i * n:(i + 1) * n] = np.tril(np.random.rand(n, n)) + 10 * np.eye(n)
        if i < N - 1:
            Lf[(i + 1) * n:(i + 2) * n, i * n:(i + 1) * n] = np.random.rand(n, n)
    A = Lf @ Lf.T
    L = np.zeros((N, n, n))
    for i in range(N):
        L[i] = A[i * n:(i + 1) * n, i * n:(i + 1) * n]
    E = np.zeros((calculate_off_diag_storage_len(N), n, n))
    for i in range(N - 1):
        E[i] = A[(i + 1) * n:(i + 2) * n, i * n:(i + 1) * n]
    np.random.seed(seed)
    b = np.random.rand(N * n).reshape(N, n, 1)
    return A, L, E, b


def dense_from_blocks(L, E):
    """Dense symmetric Psi from the diagonal blocks (lower triangle read) and
    the first N-1 E blocks (README.md:60-80)."""
    N, n, _ = L.shape
    A = np.zeros((N * n, N * n))
    for i in range(N):
        Li = np.tril(L[i])
        A[i * n:(i + 1) * n, i * n:(i + 1) * n] = Li + np.tril(Li, -1).T
        if i < N - 1:
            A[(i + 1) * n:(i + 2) * n, i * n:(i + 1) * n] = E[i]
            A[i * n:(i + 1) * n, (i + 1) * n:(i + 2) * n] = E[i].T
    return A


# ---------------------------------------------------------------------------
# batched small dense kernels (vectorised over a leading stack axis)
# ---------------------------------------------------------------------------

def _chol(A):
    """Lower Cholesky of a stack of SPD blocks; only the lower triangle is read
    (potrf.py:101-103).  numpy's LAPACK potrf reads the lower triangle too."""
    sym = np.tril(A) + np.swapaxes(np.tril(A, -1), -1, -2)
    with np.errstate(invalid="ignore"):
        try:
            return np.linalg.cholesky(sym)
        except np.linalg.LinAlgError:
            return _chol_unblocked(sym)


def _chol_unblocked(A):
    """Right-looking Cholesky that lets a non-positive pivot turn into NaN
    instead of raising, like the reference (silent NaN on non-SPD input)."""
    A = np.array(A, dtype=np.float64, copy=True)
    n = A.shape[-1]
    with np.errstate(invalid="ignore", divide="ignore"):
        for k in range(n):
            d = np.sqrt(A[..., k, k])
            A[..., k, k] = d
            A[..., k + 1:, k] /= d[..., None]
            col = A[..., k + 1:, k]
            A[..., k + 1:, k + 1:] -= col[..., :, None] * col[..., None, :]
    return np.tril(A)


def _lower_solve(Lt, X):
    """X <- Lt^{-1} X for lower-triangular Lt (tile_lower_solve_inplace)."""
    X = np.array(X, copy=True)
    n = Lt.shape[-1]
    for k in range(n):
        X[..., k, :] /= Lt[..., k, k][..., None]
        if k + 1 < n:
            X[..., k + 1:, :] -= Lt[..., k + 1:, k][..., :, None] * X[..., k, :][..., None, :]
    return X


def _upper_solve_T(Lt, X):
    """X <- Lt^{-T} X for lower-triangular Lt (tile_upper_solve_inplace on the
    transpose, trsm.py:320-335)."""
    X = np.array(X, copy=True)
    n = Lt.shape[-1]
    for k in range(n - 1, -1, -1):
        X[..., k, :] /= Lt[..., k, k][..., None]
        if k > 0:
            X[..., :k, :] -= Lt[..., k, :k][..., :, None] * X[..., k, :][..., None, :]
    return X


def _T(A):
    return np.swapaxes(A, -1, -2)


def _as4(*arrs):
    """Reshape 3-D (N, ...) arrays to a batch of one (block_tridiag_solver.py:751-756)."""
    if arrs[0].ndim == 3:
        return True, [a[None] for a in arrs]
    return False, list(arrs)


def _mask_idx(B, mask):
    if mask is None:
        return np.arange(B)
    mask = np.asarray(mask, dtype=bool)
    assert mask.shape == (B,)
    return np.nonzero(mask)[0]


# ---------------------------------------------------------------------------
# factor / forward / backward  (the ND algorithm)
# ---------------------------------------------------------------------------

def factor(L, E, mask=None, schur_operand=None):
    """In-place ND block Cholesky (create_cholesky_factor_launch, :739-923).

    Per level l (stride s), for every node i, with all lower-level updates to
    L_i already applied:
      1. L_i <- chol(L_i)                                   (potrf.py:97-119)
      2. if i+s<N:  R  <- R L_i^{-T};  L_{i+s} -= R R^T      (trsm.py:94-120, syrk.py:106-158)
      3. if i>=s:   Lc <- L_i^{-1} Lc; L_{i-s} -= Lc^T Lc    (trsm.py:202-228, syrk.py:264-315)
      4. if both:   F  -= R Lc                               (gemm.py:151-201, :345-377)
    with R = E[off+i//s] = block (i+s,i), Lc = E[off+(i-s)//s] = block (i,i-s),
    F = E[next_off+(i-s)//(2s)] = block (i+s,i-s).  The fused kernel's deferral
    of step 3's update (:88-102) is the same arithmetic reordered.
    Only tril(L) is meaningful afterwards.  Masked-out batch entries are not
    touched (README.md:214-235).
    schur_operand (optional): a function applied to R and Lc where they enter
    the Schur products of steps 2-4 (not where they are stored), to emulate a
    factor whose Schur updates take rounded operands (e.g. BF16).
    """
    op = schur_operand if schur_operand is not None else (lambda a: a)
    squeeze, (L4, E4) = _as4(L, E)
    B, N = L4.shape[:2]
    idx = _mask_idx(B, mask)
    if idx.size == 0:
        return L
    Lb = L4[idx]
    Eb = E4[idx]
    for lev in level_schedule(N):
        s, off, nxt = lev["s"], lev["off"], lev["next_off"]
        nodes = np.array(lev["nodes"])
        Lb[:, nodes] = _chol(Lb[:, nodes])
        for i in nodes:
            Li = Lb[:, i]
            has_r = i + s < N
            has_c = i >= s
            if has_r:
                r = off + i // s
                Eb[:, r] = _T(_lower_solve(Li, _T(Eb[:, r])))
            if has_c:
                c = off + (i - s) // s
                Eb[:, c] = _lower_solve(Li, Eb[:, c])
        for i in nodes:
            has_r = i + s < N
            has_c = i >= s
            if has_r:
                R = op(Eb[:, off + i // s])
                Lb[:, i + s] -= R @ _T(R)
            if has_c:
                C = op(Eb[:, off + (i - s) // s])
                Lb[:, i - s] -= _T(C) @ C
            if has_r and has_c:
                f = nxt + (i - s) // (2 * s)
                Eb[:, f] -= op(Eb[:, off + i // s]) @ op(Eb[:, off + (i - s) // s])
    L4[idx] = Lb
    E4[idx] = Eb
    return L


def round_bf16(a):
    """Rounds to the nearest BF16 value (ties to even), through FP32, as
    float64 (the rounding a BF16 MFMA operand gets from FP32 storage)."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u.astype(np.uint32)).view(np.float32).astype(np.float64)


def forward(L, E, x, mask=None):
    """In-place forward substitution y <- Lp^{-1} P b in ND order
    (create_cholesky_forward_substitution_launch, :1236-1374; kernel math
    :425-485): per level ascending, y_i <- L_i^{-1} y_i, then
    y_{i+s} -= R y_i and y_{i-s} -= Lc^T y_i."""
    squeeze, (L4, E4, x4) = _as4(L, E, x)
    B, N = L4.shape[:2]
    idx = _mask_idx(B, mask)
    if idx.size == 0:
        return x
    Lb, Eb, y = L4[idx], E4[idx], x4[idx]
    for lev in level_schedule(N):
        s, off = lev["s"], lev["off"]
        nodes = np.array(lev["nodes"])
        y[:, nodes] = _lower_solve(np.tril(Lb[:, nodes]), y[:, nodes])
        for i in nodes:
            if i + s < N:
                y[:, i + s] -= Eb[:, off + i // s] @ y[:, i]
            if i >= s:
                y[:, i - s] -= _T(Eb[:, off + (i - s) // s]) @ y[:, i]
    x4[idx] = y
    return x


def backward(L, E, x, mask=None):
    """In-place backward substitution x <- P^T Lp^{-T} y
    (create_cholesky_backward_substitution_launch, :1377-1517; kernel math
    :581-643): per level descending, x_i -= R^T x_{i+s} + Lc x_{i-s}, then
    x_i <- L_i^{-T} x_i."""
    squeeze, (L4, E4, x4) = _as4(L, E, x)
    B, N = L4.shape[:2]
    idx = _mask_idx(B, mask)
    if idx.size == 0:
        return x
    Lb, Eb, y = L4[idx], E4[idx], x4[idx]
    for lev in reversed(level_schedule(N)):
        s, off = lev["s"], lev["off"]
        nodes = np.array(lev["nodes"])
        for i in nodes:
            if i + s < N:
                y[:, i] -= _T(Eb[:, off + i // s]) @ y[:, i + s]
            if i >= s:
                y[:, i] -= Eb[:, off + (i - s) // s] @ y[:, i - s]
        y[:, nodes] = _upper_solve_T(np.tril(Lb[:, nodes]), y[:, nodes])
    x4[idx] = y
    return x


def solve(L, E, x, mask=None):
    """forward then backward (create_cholesky_solve_launch, :1520-1727)."""
    forward(L, E, x, mask)
    backward(L, E, x, mask)
    return x


def factor_and_solve(L, E, x, mask=None):
    """factor, forward, backward (create_cholesky_factor_and_solve_launch,
    :926-1233, which interleaves the forward sweep into the factor's levels;
    the arithmetic per block is the same)."""
    factor(L, E, mask)
    solve(L, E, x, mask)
    return x


# ---------------------------------------------------------------------------
# independent dense checks
# ---------------------------------------------------------------------------

def dense_nd_factor(L, E):
    """Dense oracle for the factor: Cholesky of P Psi P^T with P the ND order.

    Returns (Lp, order, L_expect, E_expect) where L_expect[i] = Lp[i,i] block and
    E_expect holds, for every level segment entry k (coupling a=(k+1)s-1,
    b=(k+2)s-1), Lp[b,a] if a is eliminated at level l (k even) else Lp[a,b]^T.
    Entries past a level's segment never exist; trailing unused E blocks stay 0.
    """
    N, n, _ = L.shape
    A = dense_from_blocks(L, E)
    order = nd_order(N)
    perm = np.concatenate([np.arange(i * n, (i + 1) * n) for i in order])
    Lp = np.linalg.cholesky(A[np.ix_(perm, perm)])
    pos = {node: p for p, node in enumerate(order)}

    def blk(r, c):
        pr, pc = pos[r], pos[c]
        return Lp[pr * n:(pr + 1) * n, pc * n:(pc + 1) * n]

    L_exp = np.stack([blk(i, i) for i in range(N)]) if N else np.zeros((0, n, n))
    E_exp = np.zeros((calculate_off_diag_storage_len(N), n, n))
    for lev in level_schedule(N):
        s, off = lev["s"], lev["off"]
        for k in range(N // s - 1):
            a, b = (k + 1) * s - 1, (k + 2) * s - 1
            if pos[a] < pos[b]:
                E_exp[off + k] = blk(b, a)
            else:
                E_exp[off + k] = blk(a, b).T
    return Lp, order, L_exp, E_exp


def block_tridiag_matvec(L, E, x):
    """Psi @ x in float64 using only tril(L) and E[:N-1]; shapes (B?,N,n,m)."""
    squeeze, (L4, E4, x4) = _as4(L, E, x)
    N = L4.shape[1]
    Lt = np.tril(L4)
    S = Lt + _T(np.tril(L4, -1))
    y = S @ x4
    if N > 1:
        Eo = E4[:, : N - 1]
        y[:, 1:] += Eo @ x4[:, :-1]
        y[:, :-1] += _T(Eo) @ x4[:, 1:]
    return y[0] if squeeze else y


def rel_residual(L, E, x, b):
    """||b - Psi x||_F / ||b||_F per batch entry (float64)."""
    r = b - block_tridiag_matvec(L, E, x)
    if r.ndim == 3:
        return np.linalg.norm(r) / np.linalg.norm(b)
    return np.linalg.norm(r.reshape(r.shape[0], -1), axis=1) / np.linalg.norm(b.reshape(b.shape[0], -1), axis=1)


# ---------------------------------------------------------------------------
# algorithmic counts used for GFLOP/s and roofline figures
# ---------------------------------------------------------------------------

def nd_flops(N: int, n: int, m: int):
    """(F_factor, F_solve) of the ND algorithm: potrf n^3/3 per node, each trsm
    and each syrk n^3, the fill-in gemm 2n^3; the solve 2n^2m per node trsm
    (forward + backward) and 2n^2m per coupling gemv in each sweep."""
    ff = 0.0
    fs = 0.0
    for lev in level_schedule(N):
        s = lev["s"]
        for i in lev["nodes"]:
            r = i + s < N
            c = i >= s
            ff += n ** 3 / 3 + (r + c) * 2 * n ** 3 + (r and c) * 2 * n ** 3
            fs += 2 * n * n * m + 4 * (r + c) * n * n * m
    return ff, fs


def nd_bytes(N: int, n: int, m: int, w: int = 8):
    """Compulsory HBM bytes: (factor, solve, factor_and_solve).
    factor: read+write tril L, read the N-1 input E blocks, write every E block.
    solve: read tril L and all E twice (two sweeps), read+write x twice.
    """
    nE = calculate_off_diag_storage_len(N)
    tri = N * n * (n + 1) // 2
    bf = w * (2 * tri + (N - 1) * n * n + nE * n * n)
    bs = w * (2 * (tri + nE * n * n) + 4 * N * n * m)
    bfs = bf + w * (tri + nE * n * n + 4 * N * n * m)
    return bf, bs, bfs
