"""BF16 Schur-complement updates (FP32 storage, FP64 diagonal blocks, BF16
MFMA products with FP32 accumulation) and the mixed solver built on them.

* The factor (socu_cholesky_factor_f32 with SOCU_PREC_BF16, through the C
  ABI) against the oracle's factor with BF16-rounded Schur operands
  (oracle.nd.factor(schur_operand=round_bf16)): the GPU factor must sit close
  to that emulation -- the remaining difference is FP32 storage and
  accumulation -- and about a BF16 rounding away from the exact FP64 factor.
  The same check on the FP32 option shows the reverse, so the two options
  cannot be confused.
* The mixed solver with precision="bfloat16" refines to rel-resid <= 1e-10
  and matches the FP64 oracle's x.
"""
import ctypes

import numpy as np
import pytest

from oracle import nd

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel())


def _gpu_factor_f32(L0, E0, precision, device):
    import torch

    from socu_amd import _lib

    lib = _lib.load()
    B, N, n, _ = L0.shape
    L = torch.tensor(L0, dtype=torch.float32, device=device)
    E = torch.tensor(E0, dtype=torch.float32, device=device)
    st = ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
    _lib.check(lib.socu_cholesky_factor_f32(L.data_ptr(), E.data_ptr(), None, B, N, n, precision, st), "factor_f32")
    torch.cuda.synchronize(device)
    return np.tril(L.double().cpu().numpy()), E.double().cpu().numpy()


@pytest.mark.parametrize("n,N", [(128, 24), (64, 33), (32, 40), (10, 17)])
def test_bf16_factor_matches_the_bf16_emulation(n, N, device):
    from socu_amd import _lib

    L0, E0, _, _, _ = nd.readme_problem(n, N, 1, alpha=16.0, seed=17)
    L0, E0 = L0[None], E0[None]
    Lx, Ex = L0.copy(), E0.copy()
    nd.factor(Lx, Ex)
    Lb, Eb = L0.copy(), E0.copy()
    nd.factor(Lb, Eb, schur_operand=nd.round_bf16)
    Lx, Lb = np.tril(Lx), np.tril(Lb)
    gap = _rel(Lb, Lx) + _rel(Eb, Ex)  # what BF16 operands change in the factor
    assert gap > 1e-5
    Lg, Eg = _gpu_factor_f32(L0, E0, _lib.PREC_BF16, device)
    d_emul = _rel(Lg, Lb) + _rel(Eg, Eb)
    d_exact = _rel(Lg, Lx) + _rel(Eg, Ex)
    print(f"n={n} N={N}: bf16 vs emulation {d_emul:.2e}, vs exact {d_exact:.2e}, emulation vs exact {gap:.2e}")
    assert d_emul < 0.2 * gap
    assert d_exact > 0.5 * gap
    # the FP32 option is close to the exact factor instead
    Lf, Ef = _gpu_factor_f32(L0, E0, _lib.PREC_FP32, device)
    assert _rel(Lf, Lx) + _rel(Ef, Ex) < 0.2 * gap


@pytest.mark.parametrize("n,N,m", [(128, 64, 1), (64, 100, 4), (32, 33, 1)])
def test_bf16_mixed_solve_refines_to_fp64(n, N, m, device):
    import torch

    from socu_amd.mixed import create_cholesky_mixed_factor_and_solve_launch

    L0, E0, b0, _, _ = nd.readme_problem(n, N, m, alpha=16.0, seed=5)
    Lr, Er, xr = L0.copy(), E0.copy(), b0.copy()
    nd.factor_and_solve(Lr, Er, xr)
    x = torch.tensor(b0, device=device)
    run = create_cholesky_mixed_factor_and_solve_launch(torch.tensor(L0, device=device), torch.tensor(E0, device=device),
                                                        x, precision="bfloat16", tol=1e-10, max_iter=30,
                                                        use_cuda_graph=True)
    run()
    rr, it = run.stats()
    xg = x.cpu().numpy()
    print(f"bf16 mixed n={n} N={N} m={m}: {it[0]} correction solves, rel-resid {rr[0]:.2e}")
    assert rr[0] <= 1e-10
    assert nd.rel_residual(L0, E0, xg, b0) <= 1e-10
    assert _rel(xg, xr) < 1e-8
