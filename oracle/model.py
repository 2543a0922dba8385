"""Forward model, NLL, gradient and MGVI metric product (TEST INFRASTRUCTURE ONLY).

Composition (per module M):
    lam_M = Readout_M(Exposure_M(linpatch_convolve(a * s, K_M)))
as in jubik/instruments/chandra/chandra_response.py:136-142 and
jubik/instruments/erosita/erosita_response.py:349-350 (the effective area
is folded into the exposure array), with
    NLL = sum_M sum_i (lam_i - d_i log lam_i)
(demos/x-ray-imaging.py:312-313).  The gradient is the hand-written
reverse chain; the reference obtains it with JAX autodiff.
"""
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import convolve as cv
from . import sky as gp
from .data import Domain


@dataclass
class Module:
    """One telescope module (or observation) on the sky grid."""
    window: np.ndarray          # (P, [E,] 2dx, 2dy) cut PSF windows
    exposure: np.ndarray        # (E, H, W) exposure (x effective area)
    keep: np.ndarray            # (E, H, W) bool, True where read out
    data: np.ndarray            # 1-D counts in C order over keep


@dataclass
class Problem:
    shape: tuple                # sky (E, H, W)
    distances: tuple            # (1, dist, dist)
    n: int
    margin: int
    pixel_area: float
    amp_ref: np.ndarray         # (G0, G1)
    amp_alpha: Optional[np.ndarray]
    zm: float
    mu_alpha: float
    dnu: np.ndarray             # (E,)
    modules: List[Module] = field(default_factory=list)

    @property
    def domain(self):
        return Domain(tuple(self.shape), tuple(self.distances))


def sky(prob, xi_ref, xi_alpha):
    return gp.gp_sky(xi_ref, xi_alpha, prob.amp_ref, prob.amp_alpha,
                     prob.zm, prob.mu_alpha, prob.dnu, crop=prob.shape[-2:])


def response(prob, s):
    """Dense lam_M (E, H, W) per module (masked pixels still present)."""
    out = []
    for mod in prob.modules:
        conv = cv.linpatch_convolve_window(prob.pixel_area * s, prob.domain,
                                           mod.window, prob.n, prob.margin)
        out.append(mod.exposure * conv)
    return out


def nll(prob, xi_ref, xi_alpha=None):
    s = sky(prob, xi_ref, xi_alpha)
    total = 0.0
    for mod, lam in zip(prob.modules, response(prob, s)):
        v = lam[mod.keep]
        total += float(np.sum(v) - np.sum(mod.data * np.log(v)))
    return total


def _log_sky_cotangent(prob, s, cot):
    """Cotangent of log s for the dense per-module cotangents ``cot`` of lam."""
    xbar = np.zeros(s.shape)
    for mod, c in zip(prob.modules, cot):
        xbar += cv.linpatch_adjoint_window(mod.exposure * c, prob.domain,
                                           mod.window, prob.n, prob.margin)
    return s * (prob.pixel_area * xbar)


def _backward(prob, s, cot):
    """J^T applied to the dense per-module cotangents ``cot`` of lam."""
    g_ref, g_alpha, _, _ = gp.gp_adjoint(
        _log_sky_cotangent(prob, s, cot), prob.amp_ref, prob.amp_alpha, prob.dnu,
        grid_shape=prob.amp_ref.shape)
    return g_ref, g_alpha


def _nll_and_cotangents(prob, s):
    total = 0.0
    cot = []
    for mod, lam in zip(prob.modules, response(prob, s)):
        v = lam[mod.keep]
        total += float(np.sum(v) - np.sum(mod.data * np.log(v)))
        r = np.zeros(lam.shape)
        r[mod.keep] = 1.0 - mod.data / v
        cot.append(r)
    return total, cot


def nll_and_grad(prob, xi_ref, xi_alpha=None):
    """NLL and its gradient with respect to (xi_ref, xi_alpha)."""
    s = sky(prob, xi_ref, xi_alpha)
    total, cot = _nll_and_cotangents(prob, s)
    return total, _backward(prob, s, cot)


def nll_and_grad_hyper(prob, xi_ref, xi_alpha=None):
    """NLL, its gradient with respect to (xi_ref, xi_alpha), and with respect
    to the sky model's amplitude fields and offsets.

    With s_e = exp(zm + HT(A_ref xi_ref) + (HT(A_alpha xi_alpha) + mu) dnu_e)
    (spectral_product_mf_sky.py:171-298 with the amplitude fields given) and
    u_ref = sum_e u_e, u_alpha = sum_e dnu_e u_e for the cotangent u of log s:
    d/dA_ref = xi_ref * HT(u_ref), d/dA_alpha = xi_alpha * HT(u_alpha),
    d/dzm = sum u_ref, d/dmu = sum u_alpha (HT is symmetric).  Returns
    (nll, g_ref, g_alpha, g_amp_ref, g_amp_alpha, g_zm, g_mu); the alpha
    entries are None for a single-energy sky.
    """
    s = sky(prob, xi_ref, xi_alpha)
    total, cot = _nll_and_cotangents(prob, s)
    u_log = _log_sky_cotangent(prob, s, cot)
    grid = prob.amp_ref.shape
    full = np.zeros(u_log.shape[:-2] + tuple(grid))
    full[..., :u_log.shape[-2], :u_log.shape[-1]] = u_log
    u_ref = full.sum(axis=0)
    h_ref = gp.hartley(u_ref)
    out = [total, prob.amp_ref * h_ref, None, np.asarray(xi_ref) * h_ref, None,
           float(u_ref.sum()), None]
    if prob.amp_alpha is not None:
        u_alpha = np.tensordot(np.asarray(prob.dnu, dtype=np.float64), full, axes=(0, 0))
        h_alpha = gp.hartley(u_alpha)
        out[2] = prob.amp_alpha * h_alpha
        out[4] = np.asarray(xi_alpha) * h_alpha
        out[6] = float(u_alpha.sum())
    return tuple(out)


def jvp(prob, xi_ref, xi_alpha, v_ref, v_alpha):
    """Dense d lam_M for the tangent (v_ref, v_alpha)."""
    s = sky(prob, xi_ref, xi_alpha)
    G0, G1 = prob.amp_ref.shape
    dlog = gp.hartley(prob.amp_ref * v_ref)[None]
    if v_alpha is not None:
        dlog = dlog + gp.hartley(prob.amp_alpha * v_alpha)[None] * \
            np.asarray(prob.dnu)[:, None, None]
    H, W = prob.shape[-2:]
    ds = s * dlog[:, :H, :W]
    return response(prob, ds), s


def metric_product(prob, xi_ref, xi_alpha, v_ref, v_alpha=None):
    """MGVI metric J^T diag(1/lam) J v at the position xi."""
    dlam, s = jvp(prob, xi_ref, xi_alpha, v_ref, v_alpha)
    lam = response(prob, s)
    cot = []
    for mod, dl, lm in zip(prob.modules, dlam, lam):
        c = np.zeros(lm.shape)
        c[mod.keep] = dl[mod.keep] / lm[mod.keep]
        cot.append(c)
    return _backward(prob, s, cot)
