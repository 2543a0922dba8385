"""Pins of the oracle's geoVI residual (oracle/geovi.py) by what its
definition implies: the cross operator is the MGVI metric at x = p, K and
K^T are adjoint, the residual vanishes to second order at the linear sample
of its metric sample, and (I + K)^T r is the gradient of 1/2 |r|^2."""
import numpy as np
import pytest

from oracle import geovi as og
from oracle import model as om
from tests.problems import make_problem


@pytest.fixture(scope="module")
def setup():
    prob, (pr, pa) = make_problem(H=32, n=4, margin=4, E=2, M=1, seed=111, expo_range=(5.0, 20.0))
    rng = np.random.default_rng(5)
    p = (pr, pa)
    x = (pr + 0.2 * rng.normal(size=pr.shape), pa + 0.2 * rng.normal(size=pa.shape))
    v = (rng.normal(size=pr.shape), rng.normal(size=pa.shape))
    w = (rng.normal(size=pr.shape), rng.normal(size=pa.shape))
    return prob, p, x, v, w


def _dot(a, b):
    return sum(float(np.sum(u * t)) for u, t in zip(a, b) if u is not None)


def test_cross_operator_is_the_metric_at_the_base_point(setup):
    prob, p, _, v, _ = setup
    k = og.cross(prob, p, p, v, to_p=True)
    kt = og.cross(prob, p, p, v, to_p=False)
    mr, ma = om.metric_product(prob, p[0], p[1], v[0], v[1])
    for got, want in ((k, (mr, ma)), (kt, (mr, ma))):
        np.testing.assert_allclose(got[0], want[0], rtol=1e-10, atol=1e-12 * np.abs(want[0]).max())
        np.testing.assert_allclose(got[1], want[1], rtol=1e-10, atol=1e-12 * np.abs(want[1]).max())


def test_cross_operators_are_adjoint(setup):
    prob, p, x, v, w = setup
    lhs = _dot(w, og.cross(prob, p, x, v, to_p=True))
    rhs = _dot(og.cross(prob, p, x, w, to_p=False), v)
    assert lhs == pytest.approx(rhs, rel=1e-11)


def test_residual_vanishes_to_second_order_at_the_linear_sample(setup):
    prob, p, _, v, _ = setup
    res = []
    for eps in (1e-2, 5e-3, 2.5e-3):
        e = (eps * v[0], eps * v[1])
        me = om.metric_product(prob, p[0], p[1], e[0], e[1])
        m = (e[0] + me[0], e[1] + me[1])  # (I + M) e
        x = (p[0] + e[0], p[1] + e[1])
        r = og.residual(prob, p, m, x)
        res.append(np.sqrt(2 * og.half_sq(r)))
    # second order: halving eps quarters the residual
    assert res[0] / res[1] == pytest.approx(4.0, rel=0.05)
    assert res[1] / res[2] == pytest.approx(4.0, rel=0.05)


def test_gauss_newton_gradient_by_central_differences(setup):
    prob, p, x, v, w = setup
    m = (0.3 * w[0], 0.3 * w[1])
    g = og.gradient(prob, p, m, x)
    h = 1e-5
    xp = (x[0] + h * v[0], x[1] + h * v[1])
    xm = (x[0] - h * v[0], x[1] - h * v[1])
    fd = (og.half_sq(og.residual(prob, p, m, xp)) - og.half_sq(og.residual(prob, p, m, xm))) / (2 * h)
    assert _dot(g, v) == pytest.approx(fd, rel=1e-6)
