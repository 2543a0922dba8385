"""The likelihood gradient chained to the sky's scalar latents (spatial and
spectral-index fluctuations, spectral-index mean, zero mode) on the GPU,
against the oracle's chain rule, which tests/test_oracle_priors.py pins by
central differences."""
import numpy as np
import pytest
import torch

from oracle import priors as op
from tests.problems import make_problem, powerlaw_amp, to_product

pytestmark = pytest.mark.gpu

SETTINGS = {"spatial_fluctuations": (0.8, 0.3), "spectral_index_fluctuations": (0.5, 0.2),
            "spectral_index_mean": (-1.0, 0.4), "zero_mode": (-2.0, 0.7)}


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-10), (torch.float32, 2e-5)])
@pytest.mark.parametrize("cfg", [dict(H=64, n=4, margin=8, E=3, M=1, seed=83),
                                 dict(H=64, n=4, margin=8, E=2, M=2, seed=84, N=128,
                                      pixel_area=0.37)])
def test_latent_gradients_vs_oracle(ja, cfg, dtype, tol):
    prob, xi = make_problem(**cfg)
    N = prob.amp_ref.shape[0]
    shape_ref, shape_alpha = powerlaw_amp(N), 0.5 * powerlaw_amp(N)
    latents = {"spatial_fluctuations": 0.3, "spectral_index_fluctuations": -0.4,
               "spectral_index_mean": 0.2, "zero_mode": -0.1}
    want_nll, want = op.nll_and_latent_grad(prob, shape_ref, shape_alpha, SETTINGS, latents, *xi)
    model = to_product(ja, prob, dtype)
    lat = ja.priors.SpectralSkyLatents(model.sky, shape_ref, shape_alpha, **SETTINGS)
    nll, g_r, g_a, got = ja.priors.nll_and_grad_latents(model, lat, latents, *xi)
    assert abs(nll - want_nll) <= tol * abs(want_nll)
    for k in op.KEYS:
        assert got[k] == pytest.approx(want[k], rel=tol * 10, abs=tol * 10 * abs(want[k])), k
    # the xi gradient at those hyperparameters
    p = op.problem_at(prob, shape_ref, shape_alpha, SETTINGS, latents)
    from oracle import model as om
    _, (wr, wa) = om.nll_and_grad(p, *xi)
    rel = np.linalg.norm(g_r.double().cpu().numpy() - wr) / np.linalg.norm(wr)
    assert rel <= tol


def test_latent_energy_includes_the_prior(ja):
    prob, xi = make_problem(H=32, n=2, margin=8, E=2, M=1, seed=85)
    N = prob.amp_ref.shape[0]
    model = to_product(ja, prob, torch.float64)
    lat = ja.priors.SpectralSkyLatents(model.sky, powerlaw_amp(N), 0.5 * powerlaw_amp(N),
                                       **SETTINGS)
    z = {"spatial_fluctuations": 0.5, "spectral_index_fluctuations": 0.1,
         "spectral_index_mean": -0.3, "zero_mode": 0.2}
    e, (gr, ga), gz = ja.priors.latent_energy_and_grad(model, lat, z, *xi)
    nll, g_r, g_a, g0 = ja.priors.nll_and_grad_latents(model, lat, z, *xi)
    prior = 0.5 * (np.sum(xi[0] ** 2) + np.sum(xi[1] ** 2) + sum(v ** 2 for v in z.values()))
    assert e == pytest.approx(nll + prior, rel=1e-12)
    for k in z:
        assert gz[k] == pytest.approx(g0[k] + z[k], rel=1e-12)
    assert torch.allclose(gr, g_r + torch.as_tensor(xi[0], device="cuda"), rtol=1e-12)
