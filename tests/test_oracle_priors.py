"""Pins of the oracle's latent priors (oracle/priors.py): the lognormal
transform has the requested mean and standard deviation (closed form and a
seeded Monte Carlo), and the chain-rule gradients with respect to the four
latents equal central differences of the oracle NLL."""
import numpy as np
import pytest

from oracle import priors as op
from tests.problems import make_problem, powerlaw_amp

SETTINGS = {"spatial_fluctuations": (0.8, 0.3), "spectral_index_fluctuations": (0.5, 0.2),
            "spectral_index_mean": (-1.0, 0.4), "zero_mode": (-2.0, 0.7)}


@pytest.mark.parametrize("mean,std", [(1.0, 0.5), (0.3, 0.05), (2.0, 3.0)])
def test_lognormal_moments(mean, std):
    mu, sigma = op.lognormal_moments(mean, std)
    # E[exp(mu + sigma z)] = exp(mu + sigma^2 / 2), Var = (e^{sigma^2} - 1) e^{2 mu + sigma^2}
    assert np.exp(mu + 0.5 * sigma ** 2) == pytest.approx(mean, rel=1e-14)
    assert np.sqrt((np.exp(sigma ** 2) - 1) * np.exp(2 * mu + sigma ** 2)) == pytest.approx(
        std, rel=1e-13)
    z = np.random.default_rng(0).normal(size=400_000)
    v = np.exp(mu + sigma * z)
    assert v.mean() == pytest.approx(mean, rel=0.02)
    assert op.lognormal(mean, std, 0.0) == pytest.approx(np.exp(mu))


def test_normal_transform():
    assert op.normal(-1.0, 0.4, 0.0) == -1.0
    assert op.normal(-1.0, 0.4, 2.5) == pytest.approx(0.0)


@pytest.mark.parametrize("cfg", [dict(H=32, n=4, margin=4, E=3, M=1, seed=81),
                                 dict(H=32, n=4, margin=4, E=2, M=2, seed=82, N=40,
                                      pixel_area=0.37)])
def test_latent_gradients_finite_differences(cfg):
    prob, xi = make_problem(**cfg)
    N = prob.amp_ref.shape[0]
    shape_ref = powerlaw_amp(N)
    shape_alpha = 0.5 * powerlaw_amp(N)
    latents = {"spatial_fluctuations": 0.3, "spectral_index_fluctuations": -0.4,
               "spectral_index_mean": 0.2, "zero_mode": -0.1}
    _, grads = op.nll_and_latent_grad(prob, shape_ref, shape_alpha, SETTINGS, latents, *xi)
    h = 1e-6
    for k in op.KEYS:
        lp, lm = dict(latents), dict(latents)
        lp[k] += h
        lm[k] -= h
        fd = (op.nll(prob, shape_ref, shape_alpha, SETTINGS, lp, *xi) -
              op.nll(prob, shape_ref, shape_alpha, SETTINGS, lm, *xi)) / (2 * h)
        assert fd == pytest.approx(grads[k], rel=2e-6, abs=1e-6 * abs(fd) + 1e-9), k
