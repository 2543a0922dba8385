"""The oracle against the fixtures and assertions of J-UBIK's own tests.

Each test restates one reference test with the same seeded numpy inputs:
test/test_convolve.py:8-55, test/test_slice_patches.py:9-66,
test/test_build_exposure_function.py:7-52 and
test/test_build_exposure_readout_function.py:8-108.  The reference cannot
be imported here (it needs jax and nifty.re), so the expectations are the
closed forms those tests assert.
"""
import numpy as np
import pytest

from oracle import convolve as cv
from oracle import response as rsp
from oracle.data import Domain


# --------------------------------------------------- test/test_convolve.py
def test_bilinear_weights_even_shape_properties():
    w = cv.bilinear_weights((4, 4))
    assert w.shape == (4, 4)
    assert np.isfinite(w).all()
    assert np.all(w >= 0)
    np.testing.assert_allclose(w, w[::-1, :])
    np.testing.assert_allclose(w, w[:, ::-1])
    # closed form: b = [0, 1, 1, 0] -> outer product
    np.testing.assert_array_equal(w, np.outer([0, 1, 1, 0], [0, 1, 1, 0]))


def test_bilinear_weights_odd_shape_raises():
    with pytest.raises(ValueError):
        cv.bilinear_weights((3, 3))


def test_integrate_constant_field_exact():
    res = cv.integrate(np.ones((3, 4)), Domain((3, 4), (0.5, 0.25)), axes=(0, 1))
    assert res == pytest.approx(3 * 4 * 0.5 * 0.25)


def test_convolve_identity_kernel_returns_input_for_unit_cell_volume():
    signal = np.random.default_rng(0).normal(size=(4, 4))
    kernel = np.zeros((4, 4))
    kernel[0, 0] = 1.0
    res = cv.convolve(kernel, signal, Domain((4, 4), (1.0, 1.0)), axes=(-2, -1))
    np.testing.assert_allclose(res, signal, rtol=1e-10, atol=1e-10)


def test_convolve_broadcasting_branch_prints_and_preserves_channel_shape(capsys):
    kernel = np.zeros((4, 4))
    kernel[0, 0] = 1.0
    signal = np.stack([np.ones((4, 4)), 2.0 * np.ones((4, 4))])
    res = cv.convolve(kernel, signal, Domain((4, 4), (1.0, 1.0)), axes=(-2, -1))
    assert "Dimension Inconsistency. Broadcasting PSFs" in capsys.readouterr().out
    assert res.shape == signal.shape
    np.testing.assert_allclose(res, signal, rtol=1e-10, atol=1e-10)


def test_convolve_matches_direct_circular_sum():
    """FFT convolution against the textbook circular sum on 6 x 5."""
    from oracle.brute import circular_convolve_loops
    rng = np.random.default_rng(3)
    k, z = rng.normal(size=(6, 5)), rng.normal(size=(6, 5))
    res = cv.convolve(k, z, Domain((6, 5), (1.0, 1.0)), axes=(-2, -1))
    np.testing.assert_allclose(res, np.array(circular_convolve_loops(k.tolist(), z.tolist())),
                               atol=1e-12)


# --------------------------------------------- test/test_slice_patches.py
SHAPE, NP, OVERLAP = (36, 36), 16, 2


def _f(a):
    return cv.slice_patches(a, SHAPE, NP, OVERLAP)


def test_slice_patches():
    x = np.arange(SHAPE[0] * SHAPE[1], dtype=float).reshape(SHAPE)
    s = _f(x)
    dx = int((SHAPE[-2] - 2 * OVERLAP) / NP)
    dy = int((SHAPE[-1] - 2 * OVERLAP) / NP)
    assert s.shape == (NP ** 2, 2 * dx + 2 * OVERLAP, 2 * dy + 2 * OVERLAP)
    assert np.isfinite(s).all()
    # patch p = iy * n + ix starts at (ix * dx, iy * dy) of the zero-padded
    # array (pad dx // 2): patch 1 is the second window down axis -2
    xp = np.pad(x, 1)
    np.testing.assert_array_equal(s[1], xp[2:10, 0:8])
    np.testing.assert_array_equal(s[NP], xp[0:8, 2:10])


def test_slice_patches_linearity():
    rng = np.random.default_rng(0)
    a, b = rng.normal(size=SHAPE), rng.normal(size=SHAPE)
    np.testing.assert_allclose(_f(a + b), _f(a) + _f(b))


def test_adjointness():
    rng = np.random.default_rng(1)
    a = rng.random(SHAPE)
    res = _f(a)
    b = rng.random(res.shape)
    lhs = np.vdot(b, _f(a))
    rhs = np.vdot(cv.slice_patches_transpose(b, SHAPE, NP, OVERLAP), a)
    np.testing.assert_allclose(lhs, rhs)


def test_slice_patches_transpose_is_the_matrix_transpose():
    """Build the slicing matrix column by column on a small grid."""
    shape, n, m = (10, 10), 3, 1
    cols = []
    for i in range(100):
        e = np.zeros(100)
        e[i] = 1.0
        cols.append(cv.slice_patches(e.reshape(shape), shape, n, m).ravel())
    A = np.stack(cols, axis=1)
    y = np.random.default_rng(2).normal(size=A.shape[0])
    got = cv.slice_patches_transpose(
        y.reshape(cv.slice_patches(np.zeros(shape), shape, n, m).shape), shape, n, m)
    np.testing.assert_allclose(got.ravel(), A.T @ y, atol=1e-12)


# ------------------------------------- test/test_build_exposure_function.py
def _exposures():
    return np.random.default_rng(0).uniform(0.0, 3e3, size=3 * 100 ** 2).reshape((3, 100, 100))


def _x():
    return np.linspace(0.25, 1.75, num=100 * 100).reshape((100, 100))


def test_build_exposure_applies_threshold_and_multiplies_input():
    exposures, x = _exposures(), _x()
    result = rsp.build_exposure_function(exposures.copy(), 500)(x)
    expected = exposures.copy()
    expected[expected < 500] = 0
    expected = expected * x
    assert result.shape == exposures.shape
    assert np.isfinite(result).all() and (result >= 0).all()
    np.testing.assert_allclose(result, expected)


def test_build_exposure_without_cut_only_multiplies():
    exposures, x = _exposures(), _x()
    np.testing.assert_allclose(rsp.build_exposure_function(exposures.copy(), None)(x),
                               exposures * x)


def test_build_exposure_negative_cut_raises():
    with pytest.raises(ValueError):
        rsp.build_exposure_function(_exposures(), -1)


def test_build_exposure_cut_mutates_callers_array():
    e = _exposures()
    rsp.build_exposure_function(e, 500)
    assert (e[e > 0] >= 500).all()


# ------------------------------- test/test_build_exposure_readout_function.py
KEYS = ('tm_1', 'tm_2')


def _ro_exposures():
    return np.random.default_rng(0).uniform(0.0, 3e3, size=2 * 24 ** 2 * 3).reshape(2, 3, 24, 24)


def _ro_x():
    return np.arange(2 * 3 * 24 * 24, dtype=float).reshape((2, 3, 24, 24)) + 0.5


def test_build_readout_function():
    exposures, x = _ro_exposures(), _ro_x()
    flags = exposures.copy()
    mask = flags < 20
    result = rsp.build_readout_function(flags, 20, KEYS)(x)
    assert tuple(result.keys()) == KEYS
    for i, key in enumerate(KEYS):
        assert result[key].ndim == 1
        assert result[key].size == np.count_nonzero(~mask[i])
        np.testing.assert_array_equal(result[key], x[i][~mask[i]])


def test_build_readout_function_wrong_input_shape():
    ro = rsp.build_readout_function(_ro_exposures(), 20, KEYS)
    with pytest.raises(ValueError):
        ro(_ro_x()[0])


def test_build_readout_function_negative_exposure_cut():
    with pytest.raises(ValueError):
        rsp.build_readout_function(_ro_exposures(), -1, KEYS)


def test_build_readout_function_keys_length_mismatch():
    with pytest.raises(ValueError):
        rsp.build_readout_function(_ro_exposures(), 20, ("tm_1",))


def test_build_readout_function_with_none_threshold_reads_all_elements():
    x = _ro_x()
    result = rsp.build_readout_function(_ro_exposures(), None, KEYS)(x)
    for i, key in enumerate(KEYS):
        np.testing.assert_array_equal(result[key], x[i].ravel())


def test_build_readout_function_with_none_keys():
    single = np.expand_dims(_ro_exposures()[0], axis=0)
    sky = np.arange(3 * 24 * 24, dtype=float).reshape((1, 3, 24, 24)) + 1.0
    mask = single < 20
    result = rsp.build_readout_function(single.copy(), 20, None)(sky)
    assert tuple(result.keys()) == ("masked input",)
    np.testing.assert_array_equal(result["masked input"], sky[0][~mask[0]])


def test_build_readout_function_wrong_flags_shape_raises():
    ro = rsp.build_readout_function(np.ones((1, 4, 4)), threshold=None, keys=KEYS[:1])
    with pytest.raises(ValueError, match="flags should have shape"):
        ro(np.ones((1, 2, 4, 4)))
