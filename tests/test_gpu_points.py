"""Point sources on the GPU (jubik_model_set_point_sources,
jubik_model_grad_points, jubik_invgamma_prior) against the oracle
(oracle/points.py, pinned by tests/test_oracle_points.py)."""
import numpy as np
import pytest
import torch

from oracle import points as op
from tests.problems import make_problem, to_product

pytestmark = pytest.mark.gpu

PS = dict(a=1.995, q=0.8, alpha_mean=-1.5, alpha_std=0.4)


def _rel(a, b):
    a = a.double().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


@pytest.mark.parametrize("a,q", [(1.9952748894654722, 1.6281324151282162e-09), (1.0001, 3e-4),
                                 (0.5, 2.0), (10.0, 1.0)])
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 2e-6)])
def test_invgamma_prior_vs_oracle(ja, a, q, dtype, tol):
    xi = np.concatenate([np.linspace(-8.0, 8.0, 321), [-20.0, -12.0, 12.0, 20.0]])
    prior = ja.priors.invgamma_prior(a, q)
    logi, dlog = prior.log_and_dlog(torch.as_tensor(xi, dtype=dtype, device="cuda"))
    want_log = np.log(op.invgamma(xi, a, q))
    want_dlog = op.dlog_invgamma(xi, a, q)
    got_log = logi.double().cpu().numpy()
    assert np.max(np.abs(got_log - want_log) / np.maximum(1.0, np.abs(want_log))) <= tol
    np.testing.assert_allclose(dlog.double().cpu().numpy(), want_dlog, rtol=max(tol, 1e-11) * 10)
    # the transform itself
    np.testing.assert_allclose(prior(xi[:321]), op.invgamma(xi[:321], a, q), rtol=tol * 20)


CFGS = [dict(H=64, n=4, margin=8, E=2, M=1, seed=101),
        dict(H=64, n=4, margin=8, E=1, M=2, seed=102, pixel_area=0.37),
        dict(H=64, n=4, margin=8, E=3, M=2, seed=103, N=128)]


def _latents(prob, seed):
    rng = np.random.default_rng(seed)
    H = prob.shape[-1]
    xps = rng.normal(size=(H, H))
    xpa = rng.normal(size=(H, H)) if prob.amp_alpha is not None else None
    return xps, xpa


@pytest.mark.parametrize("cfg", CFGS)
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-10), (torch.float32, 1e-5)])
def test_point_source_gradients_vs_oracle(ja, cfg, dtype, tol):
    prob, (xr, xa) = make_problem(**cfg)
    xps, xpa = _latents(prob, cfg["seed"])
    want_nll, (wr, wa, wps, wpa) = op.nll_and_grad(prob, xr, xa, xps, xpa, PS)
    model = to_product(ja, prob, dtype)
    model.set_point_sources(PS["a"], PS["q"], PS["alpha_mean"], PS["alpha_std"])
    nll, gr, ga, gps, gpa = model.nll_and_grad_points(xr, xa, xps, xpa)
    assert abs(nll - want_nll) <= tol * abs(want_nll)
    assert _rel(gr, wr) <= tol
    assert _rel(gps, wps) <= tol
    if wa is not None:
        assert _rel(ga, wa) <= tol
        assert _rel(gpa, wpa) <= tol
    else:
        assert ga is None and gpa is None


def test_point_source_metric_at_the_total_sky(ja):
    prob, (xr, xa) = make_problem(**CFGS[0])
    xps, xpa = _latents(prob, 7)
    model = to_product(ja, prob, torch.float64)
    model.set_point_sources(PS["a"], PS["q"], PS["alpha_mean"], PS["alpha_std"])
    model.nll_and_grad_points(xr, xa, xps, xpa)
    rng = np.random.default_rng(8)
    vr, va = rng.normal(size=xr.shape), rng.normal(size=xa.shape)
    mr, ma = model.metric(vr, va)
    wr, wa = op.metric_product(prob, xr, xa, xps, xpa, PS, vr, va)
    assert _rel(mr, wr) <= 1e-10 and _rel(ma, wa) <= 1e-10


def test_point_source_batch_equals_single_evaluations(ja):
    prob, (xr, xa) = make_problem(**CFGS[0])
    model = to_product(ja, prob, torch.float64, max_samples=3)
    model.set_point_sources(PS["a"], PS["q"], PS["alpha_mean"], PS["alpha_std"])
    rng = np.random.default_rng(9)
    H = prob.shape[-1]
    xrs = np.stack([xr + 0.05 * rng.normal(size=xr.shape) for _ in range(3)])
    xas = np.stack([xa + 0.05 * rng.normal(size=xa.shape) for _ in range(3)])
    xps = rng.normal(size=(3, H, H))
    xpa = rng.normal(size=(3, H, H))
    nlls, gr, ga, gps, gpa = model.nll_and_grad_points(xrs, xas, xps, xpa)
    for s in range(3):
        n1, g1, a1, p1, q1 = model.nll_and_grad_points(xrs[s], xas[s], xps[s], xpa[s])
        assert nlls[s] == pytest.approx(n1, rel=1e-13)
        for got, want in ((gr[s], g1), (ga[s], a1), (gps[s], p1), (gpa[s], q1)):
            np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-11 * np.abs(want).max())


def test_point_source_errors(ja):
    prob, (xr, xa) = make_problem(**CFGS[0])
    model = to_product(ja, prob, torch.float32)
    xps, xpa = _latents(prob, 3)
    with pytest.raises(RuntimeError, match="set_point_sources"):
        model.nll_and_grad_points(xr, xa, xps, xpa)
    with pytest.raises(ValueError, match="a > 0"):
        model.set_point_sources(-1.0, 1.0)
    with pytest.raises(ValueError):
        ja.priors.invgamma_prior(0.0, 1.0)
    model.set_point_sources(PS["a"], PS["q"])
    with pytest.raises(ValueError, match="xi_ps_alpha"):
        model.nll_and_grad_points(xr, xa, xps, None)
    with pytest.raises(ValueError, match="shape"):
        model.nll_and_grad_points(xr, xa, xps[:10], xpa)
    # switching them off restores the plain model
    model.set_point_sources(None, None)
    n0, g0, _ = model.nll_and_grad(xr, xa)
    assert np.isfinite(n0)
