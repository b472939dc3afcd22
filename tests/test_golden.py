"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
from the dense ND-permuted Cholesky): the oracle must reproduce them on the
CPU, and the HIP path on the GPU.  The y stored is the forward-sweep output
(x holds y after create_cholesky_forward_substitution_launch)."""
import glob
import os

import numpy as np
import pytest

from oracle import nd

HERE = os.path.dirname(os.path.abspath(__file__))
FILES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))


def load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_fixtures_exist():
    assert len(FILES) >= 50


@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f) for f in FILES])
def test_oracle_reproduces_golden(path):
    g = load(path)
    L, E, x = g["L"].copy(), g["E"].copy(), g["b"].copy()
    nd.factor(L, E)
    np.testing.assert_allclose(np.tril(L), g["L_factor"], atol=1e-12)
    np.testing.assert_allclose(E, g["E_factor"], atol=1e-12)
    nd.forward(L, E, x)
    np.testing.assert_allclose(x, g["y"], atol=1e-12)
    nd.backward(L, E, x)
    np.testing.assert_allclose(x, g["x"], atol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f) for f in FILES])
def test_hip_reproduces_golden(path, device):
    import torch

    from socu_amd.block_tridiag_solver import (
        create_cholesky_backward_substitution_launch,
        create_cholesky_factor_launch,
        create_cholesky_forward_substitution_launch,
    )

    g = load(path)
    L = torch.tensor(g["L"], device=device)
    E = torch.tensor(g["E"], device=device)
    x = torch.tensor(g["b"], device=device)
    create_cholesky_factor_launch(L, E)()
    np.testing.assert_allclose(np.tril(L.cpu().numpy()), g["L_factor"], atol=1e-12)
    np.testing.assert_allclose(E.cpu().numpy(), g["E_factor"], atol=1e-12)
    create_cholesky_forward_substitution_launch(L, E, x)()
    np.testing.assert_allclose(x.cpu().numpy(), g["y"], atol=1e-12)
    create_cholesky_backward_substitution_launch(L, E, x)()
    np.testing.assert_allclose(x.cpu().numpy(), g["x"], atol=1e-12)
