"""Point sources added to the diffuse sky (TEST INFRASTRUCTURE ONLY).

jubik/sky_models.py:376-441 builds the point-source component of SkyModel
and adds it to the diffuse sky (sky_models.py:202, add_models); the
multi-frequency variant is jubik/sky_model/multifrequency/
spectral_product_mf_point_sources.py:99-287 (MultiFrequencyInvGammaSky).
Without spectral deviations both are, per pixel and energy,

    p_e = I(xi_ps) * exp(alpha * dnu_e),   alpha = mean + std * xi_ps_alpha,

with I = invgamma_prior(a, q)(xi_ps) (sky_models.py:386-391,
spectral_product_mf_point_sources.py:244-250) and alpha from a normal prior
(PointSourceSpectralIndex, spectral_product_mf_point_sources.py:25-98;
sky_models.py:397-407 applies the slope to the relative log energies).

nifty.re's invgamma_prior is not importable here; its published meaning is
restated: the inverse-gamma variable of shape a and scale q with the same
quantile as the standard normal xi, I = scipy.stats.invgamma.ppf(Phi(xi), a,
scale=q), computed as q / y with Q(a, y) = Phi(xi) (the upper-tail form for
xi > 0, the lower-tail one otherwise, so neither tail loses precision).  The
tests pin it against scipy.stats.invgamma (ppf and cdf) and by central
differences; whether nifty.re evaluates the map exactly or from a table is
not pinned.
"""
import numpy as np
from scipy.special import gammainccinv, gammaincinv, gammaln, ndtr

from . import convolve as cv
from . import model as om
from . import sky as gp


def invgamma_y(xi, a):
    xi = np.asarray(xi, dtype=np.float64)
    return np.where(xi <= 0.0, gammainccinv(a, ndtr(xi)), gammaincinv(a, ndtr(-xi)))


def invgamma(xi, a, q):
    """I(xi) = q / y, Q(a, y) = Phi(xi)."""
    return q / invgamma_y(xi, a)


def dlog_invgamma(xi, a, q):
    """d log I / d xi = phi(xi) Gamma(a) / (y^a e^-y)."""
    xi = np.asarray(xi, dtype=np.float64)
    y = invgamma_y(xi, a)
    return np.exp(-0.5 * xi ** 2 - 0.5 * np.log(2 * np.pi) + gammaln(a) - a * np.log(y) + y)


def point_sky(xi_ps, xi_ps_alpha, ps, dnu):
    """p (E, H, W); dnu None: no spectral part (p = I)."""
    i_ref = invgamma(xi_ps, ps["a"], ps["q"])
    if dnu is None:
        return i_ref[None]
    alpha = ps["alpha_mean"] + ps["alpha_std"] * np.asarray(xi_ps_alpha)
    return i_ref[None] * np.exp(alpha[None] * np.asarray(dnu)[:, None, None])


def _dnu(prob):
    return None if prob.amp_alpha is None else prob.dnu


def nll(prob, xi_ref, xi_alpha, xi_ps, xi_ps_alpha, ps):
    s = om.sky(prob, xi_ref, xi_alpha) + point_sky(xi_ps, xi_ps_alpha, ps, _dnu(prob))
    total = 0.0
    for mod, lam in zip(prob.modules, om.response(prob, s)):
        v = lam[mod.keep]
        total += float(np.sum(v) - np.sum(mod.data * np.log(v)))
    return total


def nll_and_grad(prob, xi_ref, xi_alpha, xi_ps, xi_ps_alpha, ps):
    """NLL and gradients (g_ref, g_alpha, g_ps, g_ps_alpha) with point sources."""
    s = om.sky(prob, xi_ref, xi_alpha)
    p = point_sky(xi_ps, xi_ps_alpha, ps, _dnu(prob))
    total, cot = om._nll_and_cotangents(prob, s + p)
    # cotangent of the total sky
    xbar = np.zeros(s.shape)
    for mod, c in zip(prob.modules, cot):
        xbar += cv.linpatch_adjoint_window(mod.exposure * c, prob.domain, mod.window, prob.n,
                                           prob.margin)
    xbar *= prob.pixel_area
    g_ref, g_alpha, _, _ = gp.gp_adjoint(s * xbar, prob.amp_ref, prob.amp_alpha, prob.dnu,
                                         grid_shape=prob.amp_ref.shape)
    px = np.sum(p * xbar, axis=0)
    g_ps = dlog_invgamma(xi_ps, ps["a"], ps["q"]) * px
    g_psa = None
    if _dnu(prob) is not None:
        g_psa = ps["alpha_std"] * np.sum(p * xbar * np.asarray(prob.dnu)[:, None, None], axis=0)
    return total, (g_ref, g_alpha, g_ps, g_psa)


def metric_product(prob, xi_ref, xi_alpha, xi_ps, xi_ps_alpha, ps, v_ref, v_alpha=None):
    """J^T diag(1/lam) J v for diffuse tangents v at lam of the total sky."""
    s = om.sky(prob, xi_ref, xi_alpha)
    p = point_sky(xi_ps, xi_ps_alpha, ps, _dnu(prob))
    dlam, _ = om.jvp(prob, xi_ref, xi_alpha, v_ref, v_alpha)
    lam = om.response(prob, s + p)
    cot = []
    for mod, dl, lm in zip(prob.modules, dlam, lam):
        c = np.zeros(lm.shape)
        c[mod.keep] = dl[mod.keep] / lm[mod.keep]
        cot.append(c)
    return om._backward(prob, s, cot)
