"""geoVI nonlinear sample updates on the device.

The eROSITA and Chandra demos run nifty.re's optimize_kl with the sample
modes "nonlinear_resample" and "nonlinear_update"
(demos/erosita_demo.py:304-316, demos/configs/eROSITA_demo.yaml:148-166,
demos/configs/chandra_demo.yaml:115; the nonlinear minimiser settings come
from jubik/minimization_parser.py:424-519).  A linear MGVI sample
x0 = p + delta (p the latent mean, delta = (I + M(p))^-1 m with m the
metric sample) is moved to the geoVI sample x that solves

    r(x) = (x - p) + J(p)^T D_p (f(lam(x)) - f(lam(p))) - m = 0,

with f(lam) = 2 sqrt(lam) the Poisson likelihood's transformation (its
Fisher metric is the identity in f) and D_p = diag(lam(p)^-1/2) on read-out
pixels.  Since f'(lam) = lam^-1/2, the linearisation of r at x = p is
(I + M(p)) (x - p) - m, whose root is the linear sample.  The update
minimises 1/2 |r|^2 by Gauss-Newton: gradient
(I + K)^T r, metric (I + K)^T (I + K), K = J(p)^T D_p D_x J(x), the Newton
system solved by the device-resident CG (XrayModel.geovi_solve), the step
halved until 1/2 |r|^2 decreases.  It stops after ``maxiter`` steps or when
the mean absolute step falls to ``xtol`` (nifty.re is not available here, so
these stopping rules are stated, not mirrored).
"""
import torch

from .model import XrayModel


def _vec(x_ref, x_alpha, dtype):
    xr = torch.as_tensor(x_ref, device="cuda").to(dtype)
    xa = None if x_alpha is None else torch.as_tensor(x_alpha, device="cuda").to(dtype)
    return xr, xa


def _half_sq(r):
    return 0.5 * sum(float((t.double() ** 2).sum()) for t in r if t is not None)


def residual(model: XrayModel, pbar, m, x):
    """r(x) for the base point pbar = (p_ref, p_alpha) (the model's base
    point), metric sample m and position x; returns (r, nll at x)."""
    nll_x = model.geovi_point(*x)
    lr, la = model.geovi_apply("lh")
    r_ref = x[0] - pbar[0] + lr - m[0]
    r_alpha = None if x[1] is None else x[1] - pbar[1] + la - m[1]
    return (r_ref, r_alpha), nll_x


def nonlinear_update(model: XrayModel, p_ref, p_alpha, m_ref, m_alpha, x0_ref, x0_alpha,
                     maxiter=25, xtol=1e-4, cg_rtol=1e-6, cg_maxiter=50, max_halvings=10):
    """Move the linear sample x0 to the geoVI sample around the mean p.
    The model's base point must be p (a gradient evaluation at p).  Returns
    (x_ref, x_alpha, history)."""
    dt = model.dtype
    pbar = _vec(p_ref, p_alpha, dt)
    m = _vec(m_ref, m_alpha, dt)
    x = _vec(x0_ref, x0_alpha, dt)
    r, _ = residual(model, pbar, m, x)
    energy = _half_sq(r)
    hist = {"energy": [energy], "steps": 0, "cg_iterations": []}
    for _ in range(maxiter):
        # the geoVI point is at x: gradient (I + K)^T r and the Newton step
        kr, ka = model.geovi_apply("KT", r[0], r[1])
        g_ref = r[0] + kr
        g_alpha = None if r[1] is None else r[1] + ka
        (d_ref, d_alpha), info = model.geovi_solve(-g_ref, None if g_alpha is None else -g_alpha,
                                                   rtol=cg_rtol, maxiter=cg_maxiter)
        hist["cg_iterations"].append(info["iterations"])
        step = 1.0
        accepted = False
        for _ in range(max_halvings + 1):
            xn = (x[0] + step * d_ref, None if x[1] is None else x[1] + step * d_alpha)
            rn, _ = residual(model, pbar, m, xn)
            en = _half_sq(rn)
            if en < energy:
                accepted = True
                break
            step *= 0.5
        if not accepted:
            # the geoVI point is left at the last trial; re-set it at x
            residual(model, pbar, m, x)
            break
        x, r, energy = xn, rn, en
        hist["energy"].append(energy)
        hist["steps"] += 1
        moved = sum(float((step * t).abs().sum()) for t in (d_ref, d_alpha) if t is not None)
        size = sum(t.numel() for t in (d_ref, d_alpha) if t is not None)
        if moved / size <= xtol:
            break
    return x[0], x[1], hist


def draw_nonlinear_samples(model: XrayModel, xi_ref, xi_alpha=None, n_samples=2, seed=0,
                           draw_rtol=1e-6, draw_absdelta=None, draw_maxiter=100,
                           **update_kwargs):
    """n_samples geoVI samples around (xi_ref, xi_alpha) ("nonlinear_resample"):
    antithetic linear samples p +- delta_k with their metric samples +- m_k
    (the linear draw's CG options draw_*), each moved by nonlinear_update
    (its options in update_kwargs).  Returns (samples_ref, samples_alpha,
    info)."""
    from .sampling import draw_linear_samples
    s_ref, s_alpha, info, (m_ref, m_alpha) = draw_linear_samples(
        model, xi_ref, xi_alpha, n_samples=n_samples, seed=seed, rtol=draw_rtol,
        absdelta=draw_absdelta, maxiter=draw_maxiter, return_metric_samples=True)
    pr, pa = _vec(xi_ref, xi_alpha, model.dtype)
    out_ref, out_alpha, hists = [], [], []
    for j in range(n_samples):
        sign = 1.0 if j % 2 == 0 else -1.0
        k = j // 2
        xr, xa, h = nonlinear_update(
            model, pr, pa, sign * m_ref[k], None if m_alpha is None else sign * m_alpha[k],
            s_ref[j], None if s_alpha is None else s_alpha[j], **update_kwargs)
        out_ref.append(xr)
        out_alpha.append(xa)
        hists.append(h)
    info = dict(info) if isinstance(info, dict) else {"linear": info}
    info["nonlinear"] = hists
    samples_alpha = None if s_alpha is None else torch.stack(out_alpha)
    return torch.stack(out_ref), samples_alpha, info
