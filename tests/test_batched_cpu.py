"""Host-side checks of the batched launch and of the problem generators /
cost model, runnable without a GPU: argument validation in Python and in the
C-ABI (all of which fails before any device work), the system-table layout,
and agreement of ``socu_amd.problems`` with the oracle's generator, residual
and cost model."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import nd
from socu_amd import _lib, problems
from socu_amd.batched import create_batched_cholesky_launch


def _sys(N, n, m=1, dtype=torch.float64):
    L = torch.zeros(N, n, n, dtype=dtype)
    E = torch.zeros(problems.calculate_off_diag_storage_len(N), n, n, dtype=dtype)
    x = torch.zeros(N, n, m, dtype=dtype)
    return L, E, x


def test_cpu_tensors_are_rejected():
    L, E, x = _sys(5, 4)
    with pytest.raises(ValueError, match="no CPU implementation"):
        create_batched_cholesky_launch([L], [E], [x])
    with pytest.raises(ValueError, match="GPU"):
        create_batched_cholesky_launch([L], [E], [x], use_cuda_graph=True)


def test_python_validation():
    L, E, x = _sys(5, 4)
    with pytest.raises(ValueError, match="unknown op"):
        create_batched_cholesky_launch([L], [E], [x], op="lu")
    with pytest.raises(ValueError, match="Unsupported dtype"):
        create_batched_cholesky_launch([L], [E], [x], dtype=torch.int32)
    with pytest.raises(ValueError, match="E blocks"):
        create_batched_cholesky_launch([L, L], [E], [x, x])
    with pytest.raises(ValueError, match="mask"):
        create_batched_cholesky_launch([L], [E], [x], mask=torch.ones(2, dtype=torch.bool))
    with pytest.raises(ValueError, match="no systems"):
        create_batched_cholesky_launch([], [], [])


def test_system_table_layout():
    assert ctypes.sizeof(_lib.BatchSystem) == 40
    assert _lib.BatchSystem.N.offset == 24 and _lib.BatchSystem.reserved.offset == 36


def _create(systems, op=_lib.OP_FACTOR_AND_SOLVE, dtype=_lib.DTYPE_F64):
    lib = _lib.load()
    table = (_lib.BatchSystem * max(1, len(systems)))(*systems)
    h = ctypes.c_void_p()
    rc = lib.socu_batched_create(ctypes.byref(h), op, table, len(systems), None, dtype, 0, 0)
    return rc, h


def test_c_abi_rejects_bad_systems_before_device_work():
    p = 0x1000  # never dereferenced: validation fails first
    rc, _ = _create([_lib.BatchSystem(p, p, p, 5, 33, 1, 0)])
    assert rc == _lib.EUNSUPPORTED and "n <= 32" in _lib.last_error()
    rc, _ = _create([_lib.BatchSystem(p, p, p, 5, 8, 33, 0)])
    assert rc == _lib.EUNSUPPORTED
    rc, _ = _create([_lib.BatchSystem(p, p, p, 0, 8, 1, 0)])
    assert rc == _lib.EINVAL and "horizon" in _lib.last_error()
    rc, _ = _create([_lib.BatchSystem(p, p, None, 5, 8, 1, 0)])
    assert rc == _lib.EINVAL and "x is NULL" in _lib.last_error()
    rc, _ = _create([_lib.BatchSystem(p, p, p, 5, 8, 1, 7)])
    assert rc == _lib.EINVAL and "reserved" in _lib.last_error()
    rc, _ = _create([_lib.BatchSystem(None, p, p, 5, 8, 1, 0)])
    assert rc == _lib.EINVAL
    rc, _ = _create([_lib.BatchSystem(p, p, p, 5, 8, 1, 0)], op=9)
    assert rc == _lib.EINVAL and "operation" in _lib.last_error()
    rc, _ = _create([_lib.BatchSystem(p, p, p, 5, 8, 1, 0)], dtype=5)
    assert rc == _lib.EINVAL and "dtype" in _lib.last_error()


def test_c_abi_empty_batch_needs_no_device():
    lib = _lib.load()
    rc, h = _create([])
    assert rc == _lib.OK
    counts = (ctypes.c_int * 3)()
    tiles = (ctypes.c_int * 3)()
    waves = (ctypes.c_int * 3)()
    assert lib.socu_batched_config(h, counts, tiles, waves) == _lib.OK
    assert list(counts) == [0, 0, 0]
    assert lib.socu_batched_execute(h, None) == _lib.OK
    assert lib.socu_batched_destroy(h) == _lib.OK
    assert lib.socu_batched_max_block_size() == 32


@pytest.mark.parametrize("n,N,m,alpha,seed", [(4, 7, 2, 10.0, 0), (32, 20, 1, 10.0, 3), (16, 1, 3, 16.0, 9)])
def test_generator_matches_oracle(n, N, m, alpha, seed):
    L, E, b = problems.readme_problem(n, N, m, alpha=alpha, seed=seed)
    L1, E1, b1, _, _ = nd.readme_problem(n, N, m, alpha=alpha, seed=seed)
    np.testing.assert_array_equal(b, b1)
    np.testing.assert_allclose(L, L1, rtol=1e-14, atol=1e-12)
    np.testing.assert_allclose(E, E1, rtol=1e-14, atol=1e-12)


def test_residual_and_cost_model_match_oracle():
    L, E, b = problems.readme_problem(8, 13, 2, seed=4)
    x = np.random.RandomState(0).randn(*b.shape)
    assert problems.rel_residual(L, E, x, b) == pytest.approx(float(nd.rel_residual(L, E, x, b)), rel=1e-12)
    for N, n, m in [(1, 4, 1), (13, 8, 2), (512, 128, 1), (200, 32, 4)]:
        assert problems.nd_flops(N, n, m) == pytest.approx(nd.nd_flops(N, n, m), rel=1e-12)
        assert problems.nd_bytes(N, n, m) == nd.nd_bytes(N, n, m)
        assert problems.calculate_off_diag_storage_len(N) == nd.calculate_off_diag_storage_len(N)


def test_mpc_batch_draw_protocol():
    batch = problems.mpc_batch(40, seed=1)
    assert [(L.shape[0], L.shape[1]) for L, _, _ in batch] == problems.mpc_shapes(40, seed=1)
    for L, E, b in batch:
        N, n = L.shape[0], L.shape[1]
        assert 10 <= N <= 200 and n in problems.MPC_BLOCK_SIZES and b.shape == (N, n, 1)
        assert E.shape[0] == problems.calculate_off_diag_storage_len(N)
    # the first system is the README draw continued on the same stream
    rs = np.random.RandomState(1)
    N = int(rs.randint(10, 201))
    n = int(rs.choice(problems.MPC_BLOCK_SIZES))
    L, E, b = problems._readme_draw(rs, n, N, 1, 10.0)
    np.testing.assert_array_equal(L, batch[0][0])


def test_mpc_batch_totals():
    """Compulsory bytes and flops of factor_and_solve over the 4096-system
    batch (shapes only; the data draws are replayed, not stored)."""
    shapes = problems.mpc_shapes(4096, seed=1)
    byt = sum(problems.nd_bytes(N, n, 1)[2] for N, n in shapes)
    fl = sum(sum(problems.nd_flops(N, n, 1)) for N, n in shapes)
    assert byt == pytest.approx(9.383e9, rel=1e-3)
    assert fl == pytest.approx(2.912e10, rel=1e-3)
