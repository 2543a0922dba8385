"""Latent priors of the spectral sky and the chain of the likelihood
gradient to them.

The sky of jubik/sky_model/multifrequency/spectral_product_mf_sky.py
(build_simple_spectral_sky, 450-626) draws its scalar hyperparameters from
standardised latents z ~ N(0, 1):

    spatial fluctuations          f_ref   = lognormal_prior(mean, std)(z)
    spectral-index fluctuations   f_alpha = lognormal_prior(mean, std)(z)
    spectral-index mean           mu      = normal_prior(mean, std)(z)
    zero mode                     zm      = normal_prior(mean, std)(z)

(scaled_excitations.py:25-34 multiplies the fluctuations into the
excitations, f * xi; spectral_product_mf_sky.py:576-582 and 612-614 build
the mean and the zero mode).  lognormal_prior(mean, std) maps z to
exp(m + s z) with s^2 = log(1 + (std/mean)^2), m = log(mean) - s^2 / 2, the
lognormal variable of that mean and standard deviation; normal_prior maps
z to mean + std z.

With the sky's amplitude fields amp = f * shape (shape = hdvol * A[pd], the
normalised amplitude) the likelihood gradient with respect to the latents
follows from the amplitude-field and offset gradients of
XrayModel.nll_and_grad_hyper:

    d/dz_f  = s * sum_k amp_k * dNLL/damp_k      (df/dz = s f)
    d/dz_mu = std_mu * dNLL/dmu,   d/dz_zm = std_zm * dNLL/dzm.

The prior's own gradient of 1/2 z^2 is z; ``latent_energy_and_grad`` adds it.
"""
import math

import numpy as np
import torch

from . import _device as dev
from . import _lib

KEYS = ("spatial_fluctuations", "spectral_index_fluctuations", "spectral_index_mean",
        "zero_mode")


class lognormal_prior:
    """z -> exp(m + s z): the lognormal variable with this mean and std."""

    def __init__(self, mean, std):
        if not (mean > 0 and std > 0):
            raise ValueError("lognormal_prior needs a positive mean and std")
        self.mean, self.std = float(mean), float(std)
        self.sigma = math.sqrt(math.log1p((self.std / self.mean) ** 2))
        self.mu = math.log(self.mean) - 0.5 * self.sigma ** 2

    def __call__(self, z):
        return math.exp(self.mu + self.sigma * float(z))

    def dlog(self, z):
        """d log f / dz."""
        return self.sigma


class normal_prior:
    """z -> mean + std z."""

    def __init__(self, mean, std):
        if not std > 0:
            raise ValueError("normal_prior needs a positive std")
        self.mean, self.std = float(mean), float(std)

    def __call__(self, z):
        return self.mean + self.std * float(z)


class invgamma_prior:
    """xi -> the inverse-gamma variable of shape a and scale q with the same
    quantile as the standard normal xi (nifty.re's invgamma_prior(a, q), the
    point-source prior of jubik/sky_models.py:386-391): I = q / y with
    Q(a, y) = Phi(xi), Q the regularised upper incomplete gamma function.
    Evaluated on the device (jubik_invgamma_prior)."""

    def __init__(self, a, q):
        if not (a > 0 and q > 0):
            raise ValueError("invgamma_prior needs a > 0 and q > 0")
        self.a, self.q = float(a), float(q)

    def log_and_dlog(self, xi):
        """(log I(xi), d log I / d xi), same shape and kind as xi."""
        dev.require_device()
        t = dev.to_device(xi, dev.float_dtype(xi)).contiguous()
        logi, dlog = torch.empty_like(t), torch.empty_like(t)
        _lib.call("jubik_invgamma_prior", dev.code(t.dtype), dev.ptr(t), int(t.numel()), self.a,
                  self.q, dev.ptr(logi), dev.ptr(dlog), dev.stream())
        return dev.like_input(logi, xi), dev.like_input(dlog, xi)

    def __call__(self, xi):
        logi, _ = self.log_and_dlog(xi)
        return torch.exp(logi) if isinstance(logi, torch.Tensor) else np.exp(logi)


class SpectralSkyLatents:
    """The scalar latents of a CorrelatedFieldSky: apply them to the sky and
    chain likelihood gradients to them.

    shape_ref / shape_alpha: the normalised amplitude fields hdvol * A[pd]
    (N x N); the sky's amplitude fields become f_ref * shape_ref and
    f_alpha * shape_alpha.  Each prior is a (mean, std) pair or a prior
    object."""

    def __init__(self, sky, shape_ref, shape_alpha=None, spatial_fluctuations=(1.0, 0.5),
                 spectral_index_fluctuations=(1.0, 0.5), spectral_index_mean=(-1.0, 0.5),
                 zero_mode=(-2.0, 1.0)):
        def build(p, kind):
            return p if callable(p) else kind(*p)

        self.sky = sky
        self.shape_ref = dev.to_device(shape_ref, sky.dtype)
        self.shape_alpha = (dev.to_device(shape_alpha, sky.dtype)
                            if sky.has_alpha and shape_alpha is not None else None)
        if sky.has_alpha and self.shape_alpha is None:
            raise ValueError("a spectral sky needs shape_alpha")
        self.priors = {"spatial_fluctuations": build(spatial_fluctuations, lognormal_prior),
                       "spectral_index_fluctuations": build(spectral_index_fluctuations,
                                                            lognormal_prior),
                       "spectral_index_mean": build(spectral_index_mean, normal_prior),
                       "zero_mode": build(zero_mode, normal_prior)}

    def keys(self):
        return KEYS if self.sky.has_alpha else ("spatial_fluctuations", "zero_mode")

    def parameters(self, latents):
        """(f_ref, f_alpha, mu, zm) of the latents (dict of floats)."""
        pr = self.priors
        f_ref = pr["spatial_fluctuations"](latents["spatial_fluctuations"])
        zm = pr["zero_mode"](latents["zero_mode"])
        if not self.sky.has_alpha:
            return f_ref, None, 0.0, zm
        return (f_ref, pr["spectral_index_fluctuations"](latents["spectral_index_fluctuations"]),
                pr["spectral_index_mean"](latents["spectral_index_mean"]), zm)

    def apply(self, latents):
        """Set the sky's amplitude fields and offsets from the latents."""
        f_ref, f_alpha, mu, zm = self.parameters(latents)
        self.sky.set_parameters(
            amp_ref=f_ref * self.shape_ref,
            amp_alpha=None if f_alpha is None else f_alpha * self.shape_alpha,
            zm=zm, mu=mu if self.sky.has_alpha else None)

    def chain(self, latents, g_amp_ref, g_amp_alpha, g_zm, g_mu):
        """Likelihood gradients with respect to the latents, given those with
        respect to the sky's amplitude fields and offsets at the amplitude
        fields the latents give (the sky must hold them: ``apply``)."""
        pr = self.priors
        s = self.sky
        dot = lambda a, b: float((a.double() * b.double()).sum())
        out = {"spatial_fluctuations": pr["spatial_fluctuations"].dlog(
                   latents["spatial_fluctuations"]) * dot(g_amp_ref, s.amp_ref),
               "zero_mode": pr["zero_mode"].std * g_zm}
        if s.has_alpha:
            out["spectral_index_fluctuations"] = pr["spectral_index_fluctuations"].dlog(
                latents["spectral_index_fluctuations"]) * dot(g_amp_alpha, s.amp_alpha)
            out["spectral_index_mean"] = pr["spectral_index_mean"].std * g_mu
        return out


def nll_and_grad_latents(model, latents_obj: SpectralSkyLatents, latents, xi_ref,
                         xi_alpha=None):
    """NLL, its gradients with respect to (xi_ref, xi_alpha) and with
    respect to the scalar latents (dict), at the hyperparameters the
    latents give (applied to the model's sky first)."""
    if latents_obj.sky is not model.sky:
        raise ValueError("the latents belong to another sky")
    latents_obj.apply(latents)
    nll, g_r, g_a, ga_r, ga_a, g_zm, g_mu = model.nll_and_grad_hyper(
        torch.as_tensor(xi_ref, device="cuda").to(model.dtype),
        None if xi_alpha is None else torch.as_tensor(xi_alpha, device="cuda").to(model.dtype))
    return nll, g_r, g_a, latents_obj.chain(latents, ga_r, ga_a, g_zm, g_mu)


def latent_energy_and_grad(model, latents_obj, latents, xi_ref, xi_alpha=None):
    """The standardised energy NLL + 1/2 (|xi|^2 + sum z^2) and its gradients
    with respect to xi and the scalar latents."""
    nll, g_r, g_a, gz = nll_and_grad_latents(model, latents_obj, latents, xi_ref, xi_alpha)
    xr = torch.as_tensor(xi_ref, device="cuda").to(model.dtype)
    xa = None if xi_alpha is None else torch.as_tensor(xi_alpha, device="cuda").to(model.dtype)
    energy = nll + 0.5 * float((xr.double() ** 2).sum())
    if xa is not None:
        energy += 0.5 * float((xa.double() ** 2).sum())
    for k in latents_obj.keys():
        energy += 0.5 * float(latents[k]) ** 2
        gz[k] = gz[k] + float(latents[k])
    return energy, (g_r + xr, None if g_a is None else g_a + xa), gz
