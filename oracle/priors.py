"""Latent priors of the spectral sky (TEST INFRASTRUCTURE ONLY).

The sky of jubik/sky_model/multifrequency/spectral_product_mf_sky.py:450-626
(build_simple_spectral_sky) draws its scalar hyperparameters from
standardised latents z ~ N(0, 1):

* fluctuations of the spatial and of the spectral-index excitations,
  lognormal_prior(mean, std) (spectral_product_utils/scaled_excitations.py:
  25-34; the excitations are fluctuations * xi, scaled_excitations.py:19-20);
* the spectral-index mean, normal_prior(mean, std)
  (spectral_product_mf_sky.py:576-582);
* the zero mode, normal_prior(mean, std) (spectral_product_mf_sky.py:612-614).

nifty.re's lognormal_prior / normal_prior are not importable here; their
published meaning is restated: a lognormal variable with the given mean and
standard deviation, exp(mu + sigma z) with sigma^2 = log(1 + (std/mean)^2)
and mu = log(mean) - sigma^2 / 2 (moment matching), and mean + std z.  The
tests pin the restatement by its moments (closed form) and the chain rule by
central differences of the oracle NLL.

The amplitude fields of oracle.model.Problem are amp = f * shape, shape =
hdvol * A[pd] (A[0] = 0) the normalised amplitude, f the fluctuations.
"""
from dataclasses import replace

import numpy as np

from . import model as om

KEYS = ("spatial_fluctuations", "spectral_index_fluctuations", "spectral_index_mean",
        "zero_mode")


def lognormal_moments(mean, std):
    sigma = np.sqrt(np.log1p((std / mean) ** 2))
    return np.log(mean) - 0.5 * sigma ** 2, sigma


def lognormal(mean, std, z):
    mu, sigma = lognormal_moments(mean, std)
    return float(np.exp(mu + sigma * z))


def normal(mean, std, z):
    return float(mean + std * z)


def sky_parameters(settings, latents):
    """(f_ref, f_alpha, mu_alpha, zm) from the latents."""
    return (lognormal(*settings["spatial_fluctuations"], latents["spatial_fluctuations"]),
            lognormal(*settings["spectral_index_fluctuations"],
                      latents["spectral_index_fluctuations"]),
            normal(*settings["spectral_index_mean"], latents["spectral_index_mean"]),
            normal(*settings["zero_mode"], latents["zero_mode"]))


def problem_at(prob, shape_ref, shape_alpha, settings, latents):
    """The oracle problem with the hyperparameters the latents give."""
    f_ref, f_alpha, mu, zm = sky_parameters(settings, latents)
    return replace(prob, amp_ref=f_ref * shape_ref,
                   amp_alpha=None if shape_alpha is None else f_alpha * shape_alpha,
                   mu_alpha=mu, zm=zm)


def nll(prob, shape_ref, shape_alpha, settings, latents, xi_ref, xi_alpha=None):
    return om.nll(problem_at(prob, shape_ref, shape_alpha, settings, latents), xi_ref, xi_alpha)


def nll_and_latent_grad(prob, shape_ref, shape_alpha, settings, latents, xi_ref, xi_alpha=None):
    """NLL and its gradient with respect to the four latents, by the chain
    rule through the amplitude-field and offset gradients:
    d/dz_f = sigma_ln * sum(amp * d/damp) (df/dz = sigma_ln f, amp = f
    shape), d/dz_mu = std_mu * d/dmu, d/dz_zm = std_zm * d/dzm."""
    p = problem_at(prob, shape_ref, shape_alpha, settings, latents)
    total, _, _, g_amp_ref, g_amp_alpha, g_zm, g_mu = om.nll_and_grad_hyper(p, xi_ref, xi_alpha)
    s_ref = lognormal_moments(*settings["spatial_fluctuations"])[1]
    s_alpha = lognormal_moments(*settings["spectral_index_fluctuations"])[1]
    grads = {"spatial_fluctuations": s_ref * float(np.sum(g_amp_ref * p.amp_ref)),
             "zero_mode": settings["zero_mode"][1] * g_zm}
    if p.amp_alpha is not None:
        grads["spectral_index_fluctuations"] = s_alpha * float(np.sum(g_amp_alpha * p.amp_alpha))
        grads["spectral_index_mean"] = settings["spectral_index_mean"][1] * g_mu
    return total, grads
