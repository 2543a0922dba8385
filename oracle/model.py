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


def _backward(prob, s, cot):
    """J^T applied to the dense per-module cotangents ``cot`` of lam."""
    xbar = np.zeros(s.shape)
    for mod, c in zip(prob.modules, cot):
        xbar += cv.linpatch_adjoint_window(mod.exposure * c, prob.domain,
                                           mod.window, prob.n, prob.margin)
    u_log = s * (prob.pixel_area * xbar)
    g_ref, g_alpha, _, _ = gp.gp_adjoint(
        u_log, prob.amp_ref, prob.amp_alpha, prob.dnu,
        grid_shape=prob.amp_ref.shape)
    return g_ref, g_alpha


def nll_and_grad(prob, xi_ref, xi_alpha=None):
    """NLL and its gradient with respect to (xi_ref, xi_alpha)."""
    s = sky(prob, xi_ref, xi_alpha)
    total = 0.0
    cot = []
    for mod, lam in zip(prob.modules, response(prob, s)):
        v = lam[mod.keep]
        total += float(np.sum(v) - np.sum(mod.data * np.log(v)))
        r = np.zeros(lam.shape)
        r[mod.keep] = 1.0 - mod.data / v
        cot.append(r)
    return total, _backward(prob, s, cot)


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
