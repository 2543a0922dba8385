"""Seeded synthetic X-ray problems shared by the tests and smoke().

Inputs follow the workload descriptions: a GP sky with a power-law spectrum
(slope -4, fluctuation std 1, offset -2), per-patch unit-sum PSF windows
(elliptical Gaussians, optionally mixed with a Moffat profile), an exposure
map with a random fraction of masked pixels, and Poisson counts drawn from
the model at a true position.  Everything is numpy ``default_rng(seed)``.
"""
import numpy as np

import __graft_entry__ as _ge
from oracle import model as om

_ja = _ge.load_package()
# The synthetic inputs come from the package (the benchmark uses the same
# generators without importing the oracle).
psf_windows = _ja.synthetic.psf_windows


def powerlaw_amp(N, slope=-4.0, fluct=1.0, distance=1.0):
    """amp = hdvol * f * A[pd], A = |k|^(slope/2), A[0] = 0, sum amp^2 = f^2."""
    return _ja.powerlaw_amplitude(N, distance=distance, slope=slope, fluctuations=fluct)


def make_problem(H=64, n=4, margin=8, E=1, M=1, seed=0, N=None, mask_frac=0.1,
                 expo_range=(1e3, 1e3), moffat=None, pixel_area=1.0, sigma=(1.0, 3.0)):
    """Return (oracle Problem, (xi_ref, xi_alpha)) at the evaluation point."""
    rng = np.random.default_rng(seed)
    N = H if N is None else N
    amp_ref = powerlaw_amp(N)
    spectral = E > 1
    amp_alpha = 0.5 * amp_ref if spectral else None
    dnu = np.log(np.linspace(1.0, 2.5, E)) - np.log(1.0) if spectral else np.zeros(1)
    dx = H // n
    w = 2 * dx
    modules = []
    xi_true = rng.normal(size=(N, N))
    xa_true = rng.normal(size=(N, N)) if spectral else None
    prob = om.Problem(shape=(E, H, H), distances=(1.0, 1.0, 1.0), n=n, margin=margin,
                      pixel_area=pixel_area, amp_ref=amp_ref, amp_alpha=amp_alpha, zm=-2.0,
                      mu_alpha=-1.0 if spectral else 0.0, dnu=dnu)
    for mod in range(M):
        win = psf_windows(rng, n, E, w, sigma=sigma, moffat=moffat,
                          rotation=mod * np.pi / max(M, 1))
        ramp = np.linspace(expo_range[0], expo_range[1], H)[None, None, :]
        expo = np.broadcast_to(ramp, (E, H, H)).copy()
        expo[rng.random(expo.shape) < mask_frac] = 0.0
        keep = expo > 0
        modules.append(om.Module(window=win, exposure=expo, keep=keep, data=np.zeros(0)))
    prob.modules = modules
    s_true = om.sky(prob, xi_true, xa_true)
    for mod, lam in zip(prob.modules, om.response(prob, s_true)):
        mod.data = rng.poisson(lam[mod.keep]).astype(np.float64)
    xi = xi_true + 0.1 * rng.normal(size=xi_true.shape)
    xa = xa_true + 0.1 * rng.normal(size=xa_true.shape) if spectral else None
    return prob, (xi, xa)


def to_product(ja, prob, dtype, max_samples=1):
    """The same problem as a jubik_amd XrayModel."""
    E, H, W = prob.shape
    N = prob.amp_ref.shape[0]
    sky = ja.CorrelatedFieldSky(N, (H, W), prob.amp_ref, prob.amp_alpha, zm=prob.zm,
                                mu=prob.mu_alpha, dnu=prob.dnu, dtype=dtype)
    M = len(prob.modules)
    psf = ja.LinpatchOperator((E, H, W), prob.n, prob.margin, n_modules=M,
                              kernel_per_energy=True, dtype=dtype)
    win = np.stack([m.window for m in prob.modules])
    psf.set_kernel(win if M > 1 else win[0], layout="window")
    expo = np.stack([m.exposure for m in prob.modules])
    counts = ja.dense_counts([m.data for m in prob.modules], np.stack([m.keep for m in prob.modules]))
    return ja.XrayModel(sky, psf, prob.pixel_area, expo, counts, max_samples=max_samples)
