"""Correlated-field GP sky with a fixed amplitude field (TEST INFRASTRUCTURE ONLY).

Restates the arithmetic of
jubik/sky_model/multifrequency/spectral_product_mf_sky.py:171-219,281-298
(single spectral behaviour, no deviations) and of the diffuse component of
jubik/sky_models.py:306-341 (exp of the field, cropped to the sky grid):

    log s_e = zm + hdvol*HT(A[pd] * f * xi_ref)
                 + (hdvol*HT(A[pd] * f_a * xi_a) + mu_a) * dnu_e

with A[0] := 0 (spectral_product_mf_sky.py:177,285).  The amplitude field
enters precomposed as ``amp = hdvol * f * A[pd]``.  HT is nifty.re's
``hartley``: Re(fftn(u)) + Im(fftn(u)).  With one energy the model is the
reference-frequency distribution (spectral_product_mf_sky.py:172-173,281-298)
and xi_a, mu_a, dnu are absent.
"""
import numpy as np
import scipy.fft as sfft

from . import convolve as _cv


def hartley(u):
    """Re(FFT) + Im(FFT) over the last two axes."""
    f = sfft.fftn(np.asarray(u, dtype=np.float64), axes=(-2, -1),
                  workers=_cv._WORKERS)
    return f.real + f.imag


def gp_log_sky(xi_ref, xi_alpha, amp_ref, amp_alpha, zm, mu_alpha, dnu):
    """log s of shape (E, G0, G1) on the full GP grid."""
    t_ref = hartley(amp_ref * xi_ref)
    if xi_alpha is None:
        return (t_ref + zm)[None]
    t_alpha = hartley(amp_alpha * xi_alpha)
    dnu = np.asarray(dnu, dtype=np.float64)[:, None, None]
    return t_ref[None] + (t_alpha[None] + mu_alpha) * dnu + zm


def gp_sky(xi_ref, xi_alpha, amp_ref, amp_alpha, zm, mu_alpha, dnu,
           crop=None):
    """s = exp(log s), cropped to the sky grid (sky_models.py:340)."""
    s = np.exp(gp_log_sky(xi_ref, xi_alpha, amp_ref, amp_alpha, zm,
                          mu_alpha, dnu))
    if crop is not None:
        s = s[:, :crop[0], :crop[1]]
    return s


def gp_adjoint(u_log, amp_ref, amp_alpha, dnu, grid_shape=None):
    """Transpose of the linear part, given the cotangent of log s.

    ``u_log`` has shape (E, H, W); when H, W are smaller than the GP grid
    it is embedded at the origin (transpose of the crop).  Returns the
    cotangents of (xi_ref, xi_alpha, zm, mu_alpha).  HT is symmetric, so
    HT^T = HT.
    """
    u_log = np.asarray(u_log, dtype=np.float64)
    if grid_shape is not None and tuple(u_log.shape[-2:]) != tuple(grid_shape):
        full = np.zeros(u_log.shape[:-2] + tuple(grid_shape))
        full[..., :u_log.shape[-2], :u_log.shape[-1]] = u_log
        u_log = full
    u_ref = u_log.sum(axis=0)
    g_ref = amp_ref * hartley(u_ref)
    if amp_alpha is None:
        return g_ref, None, float(u_ref.sum()), None
    dnu = np.asarray(dnu, dtype=np.float64)
    u_alpha = np.tensordot(dnu, u_log, axes=(0, 0))
    g_alpha = amp_alpha * hartley(u_alpha)
    return g_ref, g_alpha, float(u_ref.sum()), float(u_alpha.sum())
