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


# Rectangular arrays and lengths without their own plan (zero-padded to a
# compiled length >= 2 n - 1 and folded), against numpy's fftn in float64.
@pytest.mark.parametrize("kshape,yshape,axes", [
    ((50, 70), (50, 70), (-2, -1)),          # both axes folded (128 x 192)
    ((2, 33, 128), (2, 33, 128), (-2, -1)),   # one folded, one own length
    ((3, 100, 7), (3, 100, 7), (1, 2)),       # plot.py:513's axes [1, 2]
    ((6, 5, 9), (6, 5, 9), (0, 2)),           # two non-adjacent axes
    ((4, 37), (4, 37), (-1,)),                # one axis
    ((1, 1), (1, 1), (-2, -1)),               # single element
    ((96, 1024), (96, 1024), (-2, -1)),       # own lengths, rectangular
    ((3, 700, 40), (3, 2, 700, 40), (-2, -1)),  # broadcast branch, folded to 2048 x 96
])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_convolve_general_shapes(ja, kshape, yshape, axes, dtype, capsys):
    rng = np.random.default_rng(len(kshape) + kshape[-1])
    k = rng.normal(size=kshape).astype(dtype)
    z = rng.normal(size=yshape).astype(dtype)
    dist = tuple(0.5 + 0.25 * i for i in range(len(yshape)))
    want = cv.convolve(k.astype(np.float64), z.astype(np.float64), ODomain(yshape, dist), axes,
                       verbose=False)
    got = ja.convolve(k, z, ja.Domain(yshape, dist), axes=axes)
    assert got.shape == want.shape and got.dtype == dtype
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    assert err <= (1e-12 if dtype == np.float64 else 3e-6), err


def test_convolve_size_limit(ja):
    with pytest.raises(NotImplementedError, match="2048"):
        ja.convolve(np.ones((4, 3000)), np.ones((4, 3000)), ja.Domain((4, 3000), (1.0, 1.0)),
                    axes=(-2, -1))
    with pytest.raises(NotImplementedError):
        ja.convolve(np.ones((2, 2, 2)), np.ones((2, 2, 2)), ja.Domain((2, 2, 2), (1.0,) * 3),
                    axes=(0, 1, 2))
