"""Synthetic PSF windows for the benchmark workloads (numpy, host side).

The workloads describe per-patch point-spread functions that vary across
the field: an elliptical Gaussian whose width and ellipticity grow linearly
with the patch index, optionally mixed 50/50 with a Moffat profile, each
window normalised to unit sum.  Only the centred (2 dx, 2 dx) window that
linpatch_convolve cuts from a full-frame kernel (jubik/convolve.py:107-111)
is generated.  The sky's power-law amplitude is ``sky.powerlaw_amplitude``.
"""
import numpy as np


def psf_windows(rng, n, E, w, sigma=(1.0, 3.0), ellip=(0.0, 0.3), moffat=None,
                rotation=0.0):
    """(n * n, E, w, w) unit-sum windows; the PSF centre sits at (w/2, w/2).

    sigma and ellipticity vary linearly with the patch index p = iy * n + ix
    across the field (sigma also grows by 15 % per energy bin); the major
    axis angle is drawn uniformly from ``rng`` and offset by ``rotation``.
    With ``moffat`` = beta the profile is 0.5 Gaussian + 0.5 Moffat
    (alpha = 2 sigma), each part normalised before mixing.
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
