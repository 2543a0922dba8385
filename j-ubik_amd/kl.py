"""Minimisation of the MGVI Kullback-Leibler divergence on the device.

J-UBIK infers the sky with nifty.re's ``optimize_kl``
(demos/x-ray-imaging.py:404-430, demos/erosita_demo.py:304-316): each global
iteration draws linear (MGVI) samples around the current latent mean and
then minimises the sampled KL with Newton-CG (minimizer "M", ``xtol``,
``maxiter`` 20, demos/x-ray-imaging.py:421-425).  For standardised latents
with a N(0, 1) prior and the sample offsets delta_s fixed during a
minimisation,

    KL(xi) = 1/S sum_s [ NLL(xi + delta_s) + 1/2 |xi + delta_s|^2 ]

(up to a constant), with gradient 1/S sum_s [grad NLL(xi + delta_s) +
xi + delta_s] and metric I + 1/S sum_s J_s^T diag(1/lam_s) J_s, each
Jacobian at its own sample position.  The S gradients run in one batched
pass (XrayModel.nll_and_grad_batch), the Newton system in the library's
device-resident CG (XrayModel.kl_solve).

Stopping rules (nifty.re is not available here, so these are stated, not
mirrored): a Newton-CG minimisation stops after ``maxiter`` steps or when
the mean absolute step sum|dxi| / size(xi) falls to ``xtol``; each step
halves its length up to ``max_halvings`` times until the KL decreases.
"""
import numpy as np
import torch

from .model import XrayModel
from .sampling import draw_linear_samples


def _positions(xbar, deltas):
    return None if xbar is None else (xbar[None] + deltas).contiguous()


def kl_value_and_grad(model: XrayModel, xbar_ref, xbar_alpha, deltas_ref, deltas_alpha=None):
    """(KL value (float), (g_ref, g_alpha)) at the latent mean xbar for the
    sample offsets deltas (S, N, N); the S positions become the model's
    sample positions (for kl_solve)."""
    pr = _positions(xbar_ref, deltas_ref)
    pa = _positions(xbar_alpha, deltas_alpha)
    nll, g_r, g_a = model.nll_and_grad_batch(pr, pa)
    S = pr.shape[0]
    prior = 0.5 * float((pr.double() ** 2).sum() + (0.0 if pa is None else (pa.double() ** 2)
                                                      .sum())) / S
    value = float(np.mean(nll)) + prior
    grad_r = (g_r.double().mean(0) + pr.double().mean(0)).to(model.dtype)
    grad_a = None if pa is None else (g_a.double().mean(0) + pa.double().mean(0)).to(model.dtype)
    return value, (grad_r, grad_a)


def newton_step(model: XrayModel, xbar_ref, xbar_alpha, deltas_ref, deltas_alpha=None,
                cg_rtol=1e-6, cg_maxiter=200, check_every=10):
    """The Newton direction -M^-1 grad KL at xbar (M the KL metric) with
    the gradient and KL value: (value, grad, direction, cg_info)."""
    value, (g_r, g_a) = kl_value_and_grad(model, xbar_ref, xbar_alpha, deltas_ref, deltas_alpha)
    (d_r, d_a), info = model.kl_solve(-g_r, None if g_a is None else -g_a, rtol=cg_rtol,
                                      maxiter=cg_maxiter, check_every=check_every)
    return value, (g_r, g_a), (d_r, d_a), info


def newton_cg(model: XrayModel, xbar_ref, xbar_alpha, deltas_ref, deltas_alpha=None,
              xtol=1e-6, maxiter=20, cg_rtol=1e-6, cg_maxiter=200, max_halvings=10):
    """Minimise the sampled KL over the latent mean with Newton-CG.

    Returns (xbar_ref, xbar_alpha, history) where history lists the KL value
    before each step and the accepted step lengths."""
    xr = xbar_ref.clone()
    xa = None if xbar_alpha is None else xbar_alpha.clone()
    size = xr.numel() + (0 if xa is None else xa.numel())
    hist = {"kl": [], "step": [], "cg_iterations": []}
    for _ in range(maxiter):
        value, _, (d_r, d_a), info = newton_step(model, xr, xa, deltas_ref, deltas_alpha,
                                                 cg_rtol=cg_rtol, cg_maxiter=cg_maxiter)
        hist["kl"].append(value)
        hist["cg_iterations"].append(info["iterations"])
        t = 1.0
        for _ in range(max_halvings + 1):
            nr = xr + t * d_r
            na = None if xa is None else xa + t * d_a
            new_value, _ = kl_value_and_grad(model, nr, na, deltas_ref, deltas_alpha)
            if new_value < value:
                break
            t *= 0.5
        else:
            hist["step"].append(0.0)
            break
        hist["step"].append(t)
        mean_step = float((t * d_r).abs().double().sum() + (0.0 if d_a is None else
                                                             (t * d_a).abs().double().sum()))
        xr, xa = nr, na
        if mean_step / size <= xtol:
            break
    value, _ = kl_value_and_grad(model, xr, xa, deltas_ref, deltas_alpha)
    hist["kl"].append(value)
    return xr, xa, hist


def optimize_kl(model: XrayModel, xi_ref, xi_alpha=None, n_samples=4, n_iterations=5, seed=0,
                xtol=1e-6, maxiter=20, draw_rtol=1e-6, draw_absdelta=None, draw_maxiter=100,
                sample_mode="linear_resample", nonlinear_kwargs=None):
    """The VI loop of optimize_kl: per global iteration, draw ``n_samples``
    antithetic samples around the current mean, then minimise the KL over
    the mean with Newton-CG.  sample_mode "linear_resample" draws MGVI
    samples; "nonlinear_resample" moves each linear sample to its geoVI
    sample (geovi.nonlinear_update, options in ``nonlinear_kwargs``).
    Returns (xi_ref, xi_alpha, samples_ref, samples_alpha, history)."""
    if sample_mode not in ("linear_resample", "nonlinear_resample"):
        raise ValueError(f"unsupported sample_mode {sample_mode!r}")
    xr = torch.as_tensor(xi_ref, device="cuda").to(model.dtype)
    xa = None if xi_alpha is None else torch.as_tensor(xi_alpha, device="cuda").to(model.dtype)
    history = []
    s_ref = s_alpha = None
    for it in range(n_iterations):
        if sample_mode == "nonlinear_resample":
            from .geovi import draw_nonlinear_samples
            s_ref, s_alpha, info = draw_nonlinear_samples(
                model, xr, xa, n_samples=n_samples, seed=seed + it, draw_rtol=draw_rtol,
                draw_absdelta=draw_absdelta, draw_maxiter=draw_maxiter,
                **(nonlinear_kwargs or {}))
        else:
            s_ref, s_alpha, info = draw_linear_samples(model, xr, xa, n_samples=n_samples,
                                                       seed=seed + it, rtol=draw_rtol,
                                                       absdelta=draw_absdelta,
                                                       maxiter=draw_maxiter)
        d_ref = (s_ref - xr[None]).contiguous()
        d_alpha = None if xa is None else (s_alpha - xa[None]).contiguous()
        xr, xa, hist = newton_cg(model, xr, xa, d_ref, d_alpha, xtol=xtol, maxiter=maxiter)
        hist["draw"] = info
        history.append(hist)
    return xr, xa, s_ref, s_alpha, history
