"""Seeded synthetic X-ray problems shared by the tests, smoke() and bench.py.

Inputs follow the workload descriptions: a GP sky with a power-law spectrum
(slope -4, fluctuation std 1, offset -2), per-patch unit-sum PSF windows
(elliptical Gaussians, optionally mixed with a Moffat profile), an exposure
map with a random fraction of masked pixels, and Poisson counts drawn from
the model at a true position.  Everything is numpy ``default_rng(seed)``.
"""
import numpy as np

from oracle import model as om


def powerlaw_amp(N, slope=-4.0, fluct=1.0, distance=1.0):
    """amp = hdvol * f * A[pd], A = |k|^(slope/2), A[0] = 0, sum amp^2 = f^2."""
    k = np.fft.fftfreq(N, d=distance)
    kk = np.sqrt(k[:, None] ** 2 + k[None, :] ** 2)
    amp = np.zeros((N, N))
    nz = kk > 0
    amp[nz] = kk[nz] ** (slope / 2.0)
    return amp * fluct / np.sqrt(np.sum(amp ** 2))


def psf_windows(rng, n, E, w, sigma=(1.0, 3.0), ellip=(0.0, 0.3), moffat=None,
                rotation=0.0):
    """(P, E, w, w) windows; the PSF centre sits at index (w/2, w/2).

    sigma and ellipticity vary linearly with the patch index across the
    field; with ``moffat`` = beta the profile is 0.5 Gaussian + 0.5 Moffat
    (alpha = 2 sigma).  Each window sums to one.
    """
    P = n * n
    c = np.arange(w) - w // 2
    yy, xx = np.meshgrid(c, c, indexing="ij")
    out = np.empty((P, E, w, w))
    for p in range(P):
        t = p / max(P - 1, 1)
        for e in range(E):
            sg = (sigma[0] + (sigma[1] - sigma[0]) * t) * (1.0 + 0.15 * e)
            el = ellip[0] + (ellip[1] - ellip[0]) * t
            ang = rng.uniform(0, np.pi) + rotation
            ca, sa = np.cos(ang), np.sin(ang)
            u = ca * xx + sa * yy
            v = (-sa * xx + ca * yy) / (1.0 - el)
            r2 = (u ** 2 + v ** 2) / sg ** 2
            g = np.exp(-0.5 * r2)
            if moffat is not None:
                g = 0.5 * g / g.sum() + 0.5 * (1.0 + r2 / 4.0) ** (-moffat) / \
                    ((1.0 + r2 / 4.0) ** (-moffat)).sum()
            out[p, e] = g / g.sum()
    return out


def make_problem(H=64, n=4, margin=8, E=1, M=1, seed=0, N=None, mask_frac=0.1,
                 expo_range=(1e3, 1e3), moffat=None, pixel_area=1.0):
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
        win = psf_windows(rng, n, E, w, moffat=moffat, rotation=mod * np.pi / max(M, 1))
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
