"""Pins of the oracle's point-source sky (oracle/points.py): the inverse-gamma
prior transform against scipy.stats.invgamma (quantile and distribution
function) and central differences, the NLL gradients with respect to the
point-source latents by central differences, and the metric at lam of the
total sky by its defining quadratic form."""
import numpy as np
import pytest
from scipy.special import ndtr
from scipy.stats import invgamma as sp_invgamma

from oracle import model as om
from oracle import points as op
from tests.problems import make_problem

# the demos' prior (demos/configs/chandra_demo.yaml:54-58) and a few others
PRIORS = [(1.9952748894654722, 1.6281324151282162e-09), (1.0001, 3e-4), (0.5, 2.0), (10.0, 1.0)]


@pytest.mark.parametrize("a,q", PRIORS)
def test_invgamma_is_the_normal_quantile_map(a, q):
    xi = np.linspace(-7.5, 7.5, 61)
    got = op.invgamma(xi, a, q)
    # the quantile function below the median, the inverse survival function
    # above it (Phi(xi) rounds towards 1 in the upper tail)
    lo = xi <= 0
    np.testing.assert_allclose(got[lo], sp_invgamma.ppf(ndtr(xi[lo]), a, scale=q), rtol=1e-9)
    np.testing.assert_allclose(got[~lo], sp_invgamma.isf(ndtr(-xi[~lo]), a, scale=q), rtol=1e-9)
    # the distribution function maps it back to the normal quantile
    np.testing.assert_allclose(sp_invgamma.cdf(got[lo], a, scale=q), ndtr(xi[lo]), rtol=1e-9)
    np.testing.assert_allclose(sp_invgamma.sf(got[~lo], a, scale=q), ndtr(-xi[~lo]), rtol=1e-9)
    assert np.all(np.diff(got) > 0)


@pytest.mark.parametrize("a,q", PRIORS)
def test_dlog_invgamma_finite_differences(a, q):
    xi = np.linspace(-6.0, 6.0, 25)
    h = 1e-6
    fd = (np.log(op.invgamma(xi + h, a, q)) - np.log(op.invgamma(xi - h, a, q))) / (2 * h)
    np.testing.assert_allclose(op.dlog_invgamma(xi, a, q), fd, rtol=2e-6)


PS = dict(a=1.995, q=0.8, alpha_mean=-1.5, alpha_std=0.4)


@pytest.mark.parametrize("cfg", [dict(H=32, n=4, margin=4, E=2, M=1, seed=91),
                                 dict(H=32, n=4, margin=4, E=1, M=2, seed=92, pixel_area=0.37)])
def test_point_source_gradients_finite_differences(cfg):
    prob, (xr, xa) = make_problem(**cfg)
    rng = np.random.default_rng(cfg["seed"])
    H = prob.shape[-1]
    xps = rng.normal(size=(H, H))
    xpa = rng.normal(size=(H, H)) if cfg["E"] > 1 else None
    total, (gr, ga, gps, gpa) = op.nll_and_grad(prob, xr, xa, xps, xpa, PS)
    assert total == pytest.approx(op.nll(prob, xr, xa, xps, xpa, PS), rel=1e-14)
    # central differences with step h: truncation O(h^2), rounding about
    # |NLL| eps / h (the NLL is large: exposures of 1e3 s)
    h = 1e-4
    noise = 20 * abs(total) * 2.2e-16 / h
    for (i, j) in [(3, 5), (17, 2), (30, 30)]:
        for name, arr, g in (("xi_ps", xps, gps), ("xi_ps_alpha", xpa, gpa), ("xi_ref", xr, gr)):
            if arr is None:
                continue
            p, m = arr.copy(), arr.copy()
            p[i, j] += h
            m[i, j] -= h
            args_p = dict(xi_ref=xr, xi_alpha=xa, xi_ps=xps, xi_ps_alpha=xpa)
            args_m = dict(args_p)
            args_p[name], args_m[name] = p, m
            fd = (op.nll(prob, args_p["xi_ref"], args_p["xi_alpha"], args_p["xi_ps"],
                         args_p["xi_ps_alpha"], PS) -
                  op.nll(prob, args_m["xi_ref"], args_m["xi_alpha"], args_m["xi_ps"],
                         args_m["xi_ps_alpha"], PS)) / (2 * h)
            assert g[i, j] == pytest.approx(fd, rel=1e-5, abs=noise), (name, i, j)


def test_point_source_metric_quadratic_form():
    prob, (xr, xa) = make_problem(H=32, n=4, margin=4, E=2, M=1, seed=93)
    rng = np.random.default_rng(4)
    xps, xpa = rng.normal(size=(32, 32)), rng.normal(size=(32, 32))
    vr, va = rng.normal(size=xr.shape), rng.normal(size=xa.shape)
    mr, ma = op.metric_product(prob, xr, xa, xps, xpa, PS, vr, va)
    dlam, s = om.jvp(prob, xr, xa, vr, va)
    lam = om.response(prob, s + op.point_sky(xps, xpa, PS, prob.dnu))
    quad = sum(float(np.sum(dl[m.keep] ** 2 / lm[m.keep]))
               for m, dl, lm in zip(prob.modules, dlam, lam))
    assert float(np.sum(vr * mr) + np.sum(va * ma)) == pytest.approx(quad, rel=1e-10)
