"""Stand-alone slice_patches / convolve on the GPU against the oracle and the
reference tests' fixtures (test/test_slice_patches.py, test/test_convolve.py)."""
import numpy as np
import pytest
import torch

from oracle import convolve as cv
from oracle.data import Domain as ODomain

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,n,m", [((36, 36), 16, 2), ((3, 40, 40), 4, 3), ((30, 30), 4, 0),
                                       ((2, 28, 28), 3, 2)])
def test_slice_patches_and_transpose(ja, shape, n, m):
    rng = np.random.default_rng(1)
    x = rng.normal(size=shape)
    got = ja.slice_patches(x, shape, n, m)
    want = cv.slice_patches(x, shape, n, m)
    np.testing.assert_array_equal(got, want)
    b = rng.normal(size=want.shape)
    np.testing.assert_allclose(ja.slice_patches_transpose(b, shape, n, m),
                               cv.slice_patches_transpose(b, shape, n, m), atol=1e-12)


def test_slice_patches_reference_fixture(ja):
    x = np.arange(36 * 36, dtype=float).reshape(36, 36)
    s = ja.slice_patches(x, (36, 36), 16, 2)
    assert s.shape == (256, 8, 8) and np.isfinite(s).all()


def test_convolve_identity_and_broadcast(ja, capsys):
    signal = np.random.default_rng(0).normal(size=(4, 4))
    kernel = np.zeros((4, 4))
    kernel[0, 0] = 1.0
    res = ja.convolve(kernel, signal, ja.Domain((4, 4), (1.0, 1.0)), axes=(-2, -1))
    np.testing.assert_allclose(res, signal, rtol=1e-10, atol=1e-10)
    stacked = np.stack([np.ones((4, 4)), 2.0 * np.ones((4, 4))])
    res = ja.convolve(kernel, stacked, ja.Domain((4, 4), (1.0, 1.0)), axes=(-2, -1))
    assert "Dimension Inconsistency. Broadcasting PSFs" in capsys.readouterr().out
    np.testing.assert_allclose(res, stacked, rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("N", [8, 64, 384])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_convolve_matches_oracle(ja, N, dtype):
    rng = np.random.default_rng(N)
    k = rng.normal(size=(3, N, N)).astype(dtype)
    z = rng.normal(size=(3, 2, N, N)).astype(dtype)
    dom = ODomain((N, N), (0.5, 2.0))
    want = cv.convolve(k.astype(np.float64), z.astype(np.float64), dom, axes=(-2, -1),
                       verbose=False)
    got = ja.convolve(k, z, ja.Domain((N, N), (0.5, 2.0)), axes=(-2, -1))
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    assert err <= (1e-13 if dtype == np.float64 else 2e-6)
