"""geoVI residual and its Gauss-Newton operators (TEST INFRASTRUCTURE ONLY).

nifty.re's nonlinear sample update, configured by the reference's demos
(demos/erosita_demo.py:304-316, sample modes "nonlinear_resample" /
"nonlinear_update"), moves a linear sample to the root of

    r(x) = (x - p) + J(p)^T D_p (f(lam(x)) - f(lam(p))) - m,

f(lam) = 2 sqrt(lam) (the Poisson likelihood's transformation: its Fisher
metric is the identity in f), D = diag(lam^-1/2) on read-out pixels, m the
metric sample of the linear draw.  nifty.re is not importable here; the
residual is restated from this published definition and pinned by what it
implies: K = J(p)^T D_p D_x J(x) equals the MGVI metric at x = p, K and K^T
are adjoint, r(p + e) = O(|e|^2) when m = (I + M(p)) e, and (I + K)^T r is
the gradient of 1/2 |r|^2.  Positions and tangents are (ref, alpha) pairs.
"""
import numpy as np

from . import model as om


def _lam(prob, x):
    s = om.sky(prob, x[0], x[1])
    return s, om.response(prob, s)


def _cot(prob, vals):
    """Dense cotangents with the read-out pixels of vals (masked pixels, where
    lam can be 0, are dropped)."""
    out = []
    for mod, v in zip(prob.modules, vals):
        c = np.zeros(v.shape)
        c[mod.keep] = v[mod.keep]
        out.append(c)
    return out


def lh_part(prob, p, x):
    """J(p)^T D_p (2 sqrt lam(x) - 2 sqrt lam(p))."""
    sp, lp = _lam(prob, p)
    _, lx = _lam(prob, x)
    with np.errstate(divide="ignore", invalid="ignore"):
        cot = _cot(prob, [2.0 * (np.sqrt(a) - np.sqrt(b)) / np.sqrt(b) for a, b in zip(lx, lp)])
    return om._backward(prob, sp, cot)


def residual(prob, p, m, x):
    g = lh_part(prob, p, x)
    return tuple(None if xi is None else xi - pi + gi - mi
                 for xi, pi, gi, mi in zip(x, p, g, m))


def cross(prob, p, x, v, to_p=True):
    """K v (to_p) = J(p)^T D_p D_x J(x) v, or K^T v = J(x)^T D_x D_p J(p) v."""
    sp, lp = _lam(prob, p)
    sx, lx = _lam(prob, x)
    src = x if to_p else p
    dlam, _ = om.jvp(prob, src[0], src[1], v[0], v[1])
    with np.errstate(divide="ignore", invalid="ignore"):
        cot = _cot(prob, [d / np.sqrt(a * b) for d, a, b in zip(dlam, lp, lx)])
    return om._backward(prob, sp if to_p else sx, cot)


def half_sq(r):
    return 0.5 * sum(float(np.sum(t ** 2)) for t in r if t is not None)


def gradient(prob, p, m, x):
    """(I + K)^T r."""
    r = residual(prob, p, m, x)
    kt = cross(prob, p, x, r, to_p=False)
    return tuple(None if a is None else a + b for a, b in zip(r, kt))


def gn_metric(prob, p, x, v):
    """(K + K^T + K^T K) v (the Gauss-Newton metric minus the identity)."""
    k = cross(prob, p, x, v, to_p=True)
    kt = cross(prob, p, x, v, to_p=False)
    ktk = cross(prob, p, x, k, to_p=False)
    return tuple(None if a is None else a + b + c for a, b, c in zip(k, kt, ktk))
