"""Synthetic problem generators and the algorithmic cost model.

``readme_problem`` is the generator of the reference's README
(``README.md:100-121``): D_i = tril(randn(n,n)) + alpha I and Ec_i =
randn(n,n) drawn interleaved per i, then b = randn(N,n,m);
L[i] = D_i D_i^T + Ec_{i-1} Ec_{i-1}^T and E[i] = Ec_i D_i^T, so that the
matrix is M M^T with M lower block-bidiagonal (symmetric positive definite by
construction).  ``mpc_batch`` draws the uneven MPC batch: one RandomState
stream; per system N = randint(10, 201), n = choice(8, 12, 16, 24, 32), then
that system's README draws with m = 1.

The cost model counts algorithmic work and compulsory bytes of one ND
factor / solve (``block_tridiag_solver.py:797-813`` for the factor level
loop, ``:872-914`` for the sweeps): per node potrf n^3/3, each trsm n^3, each
syrk n^3 (one symmetric half), the fill-in gemm 2n^3; per node 2n^2 m for the
two diagonal solves plus 4 n^2 m per coupling block over both sweeps.
"""
from __future__ import annotations

import numpy as np


def calculate_recursive_iterations(N: int) -> int:
    return int(N).bit_length()


def calculate_off_diag_storage_len(N: int) -> int:
    return sum((N >> lev) - 1 for lev in range(calculate_recursive_iterations(N)))


def _readme_draw(rs, n, N, m, alpha):
    D = np.zeros((N, n, n))
    Ec = np.zeros((max(N - 1, 0), n, n))
    for i in range(N):
        D[i] = np.tril(rs.randn(n, n)) + alpha * np.eye(n)
        if i < N - 1:
            Ec[i] = rs.randn(n, n)
    b = rs.randn(N, n, m)
    L = np.matmul(D, D.transpose(0, 2, 1))
    if N > 1:
        L[1:] += np.matmul(Ec, Ec.transpose(0, 2, 1))
    E = np.zeros((calculate_off_diag_storage_len(N), n, n))
    if N > 1:
        E[: N - 1] = np.matmul(Ec, D[: N - 1].transpose(0, 2, 1))
    return L, E, b


def readme_problem(n: int, N: int, m: int = 1, alpha: float = 10.0, seed: int = 0):
    """(L, E, b) of the README generator with RandomState(seed)."""
    return _readme_draw(np.random.RandomState(seed), n, N, m, alpha)


MPC_BLOCK_SIZES = (8, 12, 16, 24, 32)


def mpc_batch(count: int = 4096, seed: int = 1, alpha: float = 10.0, m: int = 1):
    """List of (L, E, b) for `count` uneven MPC systems (see module doc)."""
    rs = np.random.RandomState(seed)
    out = []
    for _ in range(count):
        N = int(rs.randint(10, 201))
        n = int(rs.choice(MPC_BLOCK_SIZES))
        out.append(_readme_draw(rs, n, N, m, alpha))
    return out


def mpc_shapes(count: int = 4096, seed: int = 1):
    """(N, n) of each system of mpc_batch without drawing the data."""
    rs = np.random.RandomState(seed)
    shapes = []
    for _ in range(count):
        N = int(rs.randint(10, 201))
        n = int(rs.choice(MPC_BLOCK_SIZES))
        shapes.append((N, n))
        # consume the same draws as _readme_draw(m=1)
        for i in range(N):
            rs.randn(n, n)
            if i < N - 1:
                rs.randn(n, n)
        rs.randn(N, n, 1)
    return shapes


def level_nodes(N: int):
    """[(s, nodes)] of each ND level: s = 2^l, nodes s-1, 3s-1, ... < N."""
    out = []
    for lev in range(calculate_recursive_iterations(N)):
        s = 1 << lev
        out.append((s, list(range(s - 1, N, 2 * s))))
    return out


def class_flops(N: int, n: int, m: int):
    """Algorithmic flops of one factor_and_solve per kernel class."""
    f = {"potrf": 0.0, "trsm": 0.0, "schur": 0.0, "forward": 0.0, "backward": 0.0, "refine": 0.0}
    for s, nodes in level_nodes(N):
        for i in nodes:
            r = i + s < N
            c = i >= s
            f["potrf"] += n ** 3 / 3
            f["trsm"] += (r + c) * n ** 3
            f["schur"] += (r + c) * n ** 3 + (r and c) * 2 * n ** 3
            f["forward"] += n * n * m + 2 * (r + c) * n * n * m
            f["backward"] += n * n * m + 2 * (r + c) * n * n * m
    return f


def nd_flops(N: int, n: int, m: int):
    """(factor flops, solve flops)."""
    f = class_flops(N, n, m)
    return f["potrf"] + f["trsm"] + f["schur"], f["forward"] + f["backward"]


def block_matvec_flops(N: int, n: int, m: int):
    """Flops of Psi x for one system: a dense n x n block times n x m per
    diagonal block and two per coupling (E_i and E_i^T)."""
    return 2 * n * n * m * (N + 2 * max(N - 1, 0))


def nd_bytes(N: int, n: int, m: int, w: int = 8):
    """Compulsory HBM bytes (factor, solve, factor_and_solve) with w-byte
    scalars: the factor reads and writes tril L, reads the N-1 input E blocks
    and writes every E block; the solve reads tril L and all E in each of the
    two sweeps and reads and writes x in each; factor_and_solve is the factor
    plus one more read of L and E and the x traffic."""
    nE = calculate_off_diag_storage_len(N)
    tri = N * n * (n + 1) // 2
    bf = w * (2 * tri + (N - 1) * n * n + nE * n * n)
    bs = w * (2 * (tri + nE * n * n) + 4 * N * n * m)
    bfs = bf + w * (tri + nE * n * n + 4 * N * n * m)
    return bf, bs, bfs


def block_tridiag_matvec(L, E, x):
    """Psi x for one system: diagonal blocks L (N,n,n) (lower triangle used,
    mirrored), sub-diagonal blocks E[:N-1] (E[i] couples row i+1, column i)."""
    N = L.shape[0]
    Ls = np.tril(L) + np.swapaxes(np.tril(L, -1), -1, -2)
    y = np.matmul(Ls, x)
    if N > 1:
        Es = E[: N - 1]
        y[1:] += np.matmul(Es, x[:-1])
        y[:-1] += np.matmul(np.swapaxes(Es, -1, -2), x[1:])
    return y


def rel_residual(L, E, x, b):
    """||b - Psi x||_F / ||b||_F of one system (arrays shaped as above)."""
    r = b - block_tridiag_matvec(L, E, x)
    return float(np.linalg.norm(r) / np.linalg.norm(b))


def class_bytes(N: int, n: int, m: int, w: int = 4):
    """Compulsory HBM bytes of one execution of each solve-side kernel class:
    a forward (or backward) sweep reads tril L and every E block and reads and
    writes x (w-byte factor scalars); one refinement iteration reads the FP64
    matrix (tril L and E[:N-1]), x and b, writes the residual (w bytes) and
    updates x (reads the correction, reads and writes x)."""
    nE = calculate_off_diag_storage_len(N)
    tri = N * n * (n + 1) // 2
    nv = N * n * m
    sweep = w * (tri + nE * n * n + 2 * nv)
    refine = 8 * (tri + (N - 1) * n * n + 2 * nv) + w * nv + (w + 16) * nv
    return {"forward": sweep, "backward": sweep, "refine": refine}
