"""Linear (MGVI) sample draws around a latent position.

J-UBIK draws its VI samples through nifty.re's optimize_kl with
sample_mode="linear_resample" and a conjugate-gradient solve per sample
(demos/x-ray-imaging.py:404-430: draw_linear_kwargs cg_kwargs absdelta and
maxiter; demos/erosita_demo.py:304-316).  The draw in the standardised
latent space of the GP is

    b = eta1 + J^T diag(lam^-1/2) eta2,   eta1, eta2 ~ N(0, 1),
    (I + J^T diag(1/lam) J) x = b,        samples = xi +- x,

so x ~ N(0, M^-1) for the MGVI metric M.  Every operator application is one
batched pass of the HIP kernels over all S systems, and the CG recurrences
run on the device (XrayModel.metric_solve, jubik_model_metric_solve).
``conjugate_gradient`` is the same recurrence for an arbitrary operator
given as a Python callable (host-driven, one synchronisation per
iteration).
"""
import numpy as np
import torch

from .model import XrayModel


def _dot(a, b):
    """Per-system FP64 dot products of tuples of (S, ...) tensors."""
    tot = None
    for x, y in zip(a, b):
        if x is None:
            continue
        d = (x.double() * y.double()).flatten(1).sum(1)
        tot = d if tot is None else tot + d
    return tot


def _bcast(c, x):
    return c.view((-1,) + (1,) * (x.dim() - 1)).to(x.dtype)


def conjugate_gradient(apply, b, rtol=1e-6, absdelta=None, maxiter=100):
    """Solve A x = b for S systems at once (A symmetric positive definite).

    ``apply`` maps a tuple of (S, ...) tensors to A times them.  A system
    stops when its residual norm falls below rtol times that of b or, with
    ``absdelta``, when its energy 1/2 x.A.x - b.x changes by less than
    absdelta in one step (the stopping rule nifty.re's CG is given at
    demos/x-ray-imaging.py:418).  Returns (x, info) with the iteration
    counts and final relative residuals per system.
    """
    x = tuple(None if v is None else torch.zeros_like(v) for v in b)
    r = tuple(None if v is None else v.clone() for v in b)
    p = tuple(None if v is None else v.clone() for v in b)
    rr = _dot(r, r)
    bb = rr.clone()
    S = rr.shape[0]
    active = torch.ones(S, dtype=torch.bool, device=rr.device)
    active &= bb > 0
    iters = torch.zeros(S, dtype=torch.int64, device=rr.device)
    energy = torch.zeros(S, dtype=torch.float64, device=rr.device)
    for _ in range(maxiter):
        if not bool(active.any()):
            break
        Ap = apply(p)
        pAp = _dot(p, Ap)
        alpha = torch.where(active, rr / torch.where(active, pAp, torch.ones_like(pAp)),
                            torch.zeros_like(rr))
        x = tuple(None if xi is None else xi + _bcast(alpha, xi) * pi for xi, pi in zip(x, p))
        r = tuple(None if ri is None else ri - _bcast(alpha, ri) * api
                  for ri, api in zip(r, Ap))
        rr_new = _dot(r, r)
        iters += active.long()
        # energy of the iterate: 1/2 x.A.x - b.x = -1/2 x.(b + r)
        e_new = -0.5 * _dot(x, tuple(None if bi is None else bi + ri for bi, ri in zip(b, r)))
        done = rr_new <= (rtol ** 2) * bb
        if absdelta is not None:
            done |= (energy - e_new).abs() <= absdelta
        energy = torch.where(active, e_new, energy)
        beta = torch.where(active, rr_new / torch.where(rr > 0, rr, torch.ones_like(rr)),
                           torch.zeros_like(rr))
        p = tuple(None if ri is None else ri + _bcast(beta, ri) * pi for ri, pi in zip(r, p))
        rr = torch.where(active, rr_new, rr)
        active &= ~done
    info = {"iterations": iters.cpu().numpy(),
            "rel_residual": torch.sqrt(rr / torch.where(bb > 0, bb, torch.ones_like(bb)))
            .cpu().numpy()}
    return x, info


def metric_solve(model: XrayModel, b_ref, b_alpha=None, rtol=1e-6, absdelta=None, maxiter=100,
                 check_every=10):
    """x = (I + J^T diag(1/lam) J)^-1 b at the last gradient position, for
    (S, N, N) right-hand sides, by the library's device-resident CG (FP64
    dot products in fixed-order reductions, no host synchronisation between
    checks every ``check_every`` iterations); returns ((x_ref, x_alpha),
    info)."""
    return model.metric_solve(b_ref, b_alpha, rtol=rtol, absdelta=absdelta, maxiter=maxiter,
                              check_every=check_every)


def draw_linear_samples(model: XrayModel, xi_ref, xi_alpha=None, n_samples=2, seed=0,
                        rtol=1e-6, absdelta=None, maxiter=100, return_metric_samples=False):
    """n_samples antithetic MGVI samples around (xi_ref, xi_alpha).

    Returns (samples_ref (n_samples, N, N), samples_alpha or None, info);
    samples 2k and 2k + 1 are xi + x_k and xi - x_k, x_k = (I + M)^-1 m_k
    with m_k the metric samples.  With return_metric_samples the tuple ends
    with (m_ref, m_alpha) (n_samples / 2, N, N) as well (geoVI needs them).
    The base point is set by a gradient evaluation at xi; noise comes from a
    seeded torch generator on the device.
    """
    if n_samples < 2 or n_samples % 2:
        raise ValueError("n_samples must be a positive even number (antithetic pairs)")
    S = n_samples // 2
    dtype = model.dtype
    xr = torch.as_tensor(xi_ref, device="cuda").to(dtype)
    xa = None if xi_alpha is None else torch.as_tensor(xi_alpha, device="cuda").to(dtype)
    model.nll_and_grad(xr, xa)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(int(seed))
    N = model.sky.N
    eta_ref = torch.randn((S, N, N), generator=gen, device="cuda", dtype=dtype)
    eta_alpha = (torch.randn((S, N, N), generator=gen, device="cuda", dtype=dtype)
                 if xa is not None else None)
    noise = torch.randn((S,) + tuple(model.exposure.shape), generator=gen, device="cuda",
                        dtype=dtype)
    l_ref, l_alpha = model.sqrt_metric_batch(noise)
    b = (eta_ref + l_ref, None if xa is None else eta_alpha + l_alpha)
    (x_ref, x_alpha), info = metric_solve(model, b[0], b[1], rtol=rtol, absdelta=absdelta,
                                          maxiter=maxiter)
    s_ref = torch.stack([xr + x_ref, xr - x_ref], dim=1).reshape((2 * S, N, N))
    s_alpha = (torch.stack([xa + x_alpha, xa - x_alpha], dim=1).reshape((2 * S, N, N))
               if xa is not None else None)
    if return_metric_samples:
        return s_ref, s_alpha, info, b
    return s_ref, s_alpha, info
