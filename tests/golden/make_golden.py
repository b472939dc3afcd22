"""Writes the small golden fixtures under tests/golden/.

Inputs come from the README generator restated in oracle/nd.py (seeded);
expected outputs come from the dense ND-permuted Cholesky (numpy/LAPACK), not
from the code under test: tril(L) of every node, every E block of every level
segment, the forward-sweep result y and the solution x.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np
import scipy.linalg as la

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import nd  # noqa: E402

CASES = [(n, N, m) for n in (1, 3, 4, 8) for N in (1, 2, 3, 4, 7, 8, 20) for m in (1, 3)]


def golden(n, N, m, seed):
    L, E, b, _, _ = nd.readme_problem(n, N, m, alpha=10.0, seed=seed)
    Lp, order, L_exp, E_exp = nd.dense_nd_factor(L, E)
    perm = np.concatenate([np.arange(i * n, (i + 1) * n) for i in order])
    inv = np.empty_like(perm)
    inv[perm] = np.arange(N * n)
    y = la.solve_triangular(Lp, b.reshape(N * n, m)[perm], lower=True)
    x = la.solve_triangular(Lp.T, y, lower=False)
    return dict(L=L, E=E, b=b, L_factor=L_exp, E_factor=E_exp,
                y=y[inv].reshape(N, n, m), x=x[inv].reshape(N, n, m))


def main():
    for n, N, m in CASES:
        seed = 1000 * n + 10 * N + m
        path = os.path.join(HERE, f"nd_n{n}_N{N}_m{m}.npz")
        np.savez_compressed(path, seed=seed, **golden(n, N, m, seed))
    print(f"wrote {len(CASES)} fixtures to {HERE}")


if __name__ == "__main__":
    main()
