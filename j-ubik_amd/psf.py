"""eROSITA PSF preparation on the GPU: drop-in for psf_interpolator.

``psf_interpolator(domain, npatch, psf_infos)`` keeps the signature of
jubik/instruments/erosita/erosita_psf_utils.py:185 and returns the same
(P, E, H, W) array of patch PSFs rolled to the origin; every patch-energy
image is computed by libjubik_hip's interpolation kernel.  With
``layout="window"`` it returns only the centred (P, E, w, w) windows
(w = 2 H / npatch) that linpatch_convolve cuts from those frames
(convolve.py:104-111), which LinpatchOperator.set_kernel(..., "window")
takes directly, avoiding the full-frame array (1 GB at 1024^2 x 4 x 64).
"""
import ctypes

import numpy as np
import torch

from . import _device as dev
from . import _lib
from .data import is_domain


def psf_interpolator(domain, npatch, psf_infos, layout="full", dtype=torch.float64,
                     device=False):
    """Patch PSFs for each energy's ``psf_infos`` dict (keys psfs, rs,
    patch_center_ids, patch_deltas, pointing_center, as built by
    erosita_psf.py:_get_psf_infos)."""
    dev.require_device()
    if not is_domain(domain):
        raise ValueError("domain has to be an instance of jubik.Domain")
    if layout not in ("full", "window"):
        raise ValueError("layout must be 'full' or 'window'")
    H, W = int(domain.shape[-2]), int(domain.shape[-1])
    dH, dW = float(domain.distances[-2]), float(domain.distances[-1])
    npatch = int(npatch)
    for ss in (H, W):
        if ss % npatch != 0 or ss % 2 != 0:
            raise ValueError
    E = len(psf_infos)
    P = npatch * npatch
    side = (H, W) if layout == "full" else (2 * (H // npatch), 2 * (W // npatch))
    out = torch.empty((P, E) + side, dtype=dtype, device="cuda")
    for e, info in enumerate(psf_infos):
        psfs = np.asarray(info["psfs"], dtype=np.float64)
        rs = np.ascontiguousarray(np.asarray(info["rs"], dtype=np.float64).ravel())
        cid = np.ascontiguousarray(np.asarray(info["patch_center_ids"], dtype=np.float64))
        if psfs.ndim != 3 or psfs.shape[0] != rs.size or cid.shape != (rs.size, 2):
            raise ValueError
        deltas = np.ascontiguousarray(np.array(list(info["patch_deltas"]), dtype=np.float64))
        pc = np.ascontiguousarray(np.array(list(info["pointing_center"]), dtype=np.float64))
        pt = torch.as_tensor(psfs, dtype=dtype, device="cuda").contiguous()
        dp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        _lib.call("jubik_psf_interpolate", dev.code(dtype), dev.ptr(pt), rs.size, psfs.shape[1],
                  psfs.shape[2], dp(rs), dp(cid), dp(deltas), dp(pc), H, W, dH, dW, npatch, e, E,
                  0 if layout == "full" else 1, dev.ptr(out), dev.stream())
    return out if device else out.cpu().numpy()
