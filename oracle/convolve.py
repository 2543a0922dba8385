"""float64 restatement of jubik/convolve.py (TEST INFRASTRUCTURE ONLY).

The algorithm keeps the reference's order of operations:
slice -> weight -> pad -> (per call) kernel shift/cut/pad/shift/normalise
-> full complex FFT convolution -> overlap-add (transpose of slicing) -> crop.
It is the CPU baseline measured beside the GPU path and the checker for it.
"""
import os
from functools import reduce

import numpy as np
import scipy.fft as sfft

_WORKERS = max(1, len(os.sched_getaffinity(0)))


def set_workers(n):
    """Number of threads scipy.fft may use (the CPU baseline reports it)."""
    global _WORKERS
    _WORKERS = max(1, int(n))


def bilinear_weights(shape):
    """Tent weights b (x) b, jubik/convolve.py:18-25.

    ``b`` rises linearly from 0 to 1 over the first half of the window and
    falls back symmetrically; only ``shape[0]`` is used (square windows).
    """
    if shape[0] % 2 != 0:
        raise ValueError("this should happen")
    half = shape[0] // 2
    ramp = np.linspace(0.0, 1.0, half, dtype=np.float64)
    tent = np.concatenate([ramp, ramp[::-1]])
    return tent[:, None] * tent[None, :]


def patch_geometry(shape, n, margin):
    """Sizes used by slice_patches, jubik/convolve.py:43-63.

    Returns (dx, dy, hx, hy, wx, wy): the patch strides, the zero padding
    applied on each side, and the window sizes.
    """
    dx = int((shape[-2] - 2 * margin) / n)
    dy = int((shape[-1] - 2 * margin) / n)
    return dx, dy, dx // 2, dy // 2, 2 * dx + 2 * margin, 2 * dy + 2 * margin


def window_starts(n, stride, padded_len, win):
    """Start offsets of the windows along one axis.

    jax.lax.dynamic_slice clamps a start index to [0, dim - size]
    (convolve.py:55-57 relies on that when the padded axis is short).
    """
    if win > padded_len:
        raise ValueError(
            f"window of {win} does not fit the padded axis of {padded_len}")
    return [min(max(i * stride, 0), padded_len - win) for i in range(n)]


def patch_order(n):
    """(ix, iy) for flat patch index p, jubik/convolve.py:59-63.

    meshgrid(..., indexing="xy") flattened: the offset along axis -2 varies
    fastest, so p = iy * n + ix.
    """
    return [(p % n, p // n) for p in range(n * n)]


def slice_patches(x, shape, n, margin):
    """Overlapping windows of x, jubik/convolve.py:28-63.

    Output shape (n*n, *x.shape[:-2], 2dx+2m, 2dy+2m).
    """
    x = np.asarray(x)
    dx, dy, hx, hy, wx, wy = patch_geometry(shape, n, margin)
    pad = [(0, 0)] * (x.ndim - 2) + [(hx, hx), (hy, hy)]
    xp = np.pad(x, pad)
    sx = window_starts(n, dx, xp.shape[-2], wx)
    sy = window_starts(n, dy, xp.shape[-1], wy)
    out = np.empty((n * n,) + x.shape[:-2] + (wx, wy), dtype=xp.dtype)
    for p, (ix, iy) in enumerate(patch_order(n)):
        out[p] = xp[..., sx[ix]:sx[ix] + wx, sy[iy]:sy[iy] + wy]
    return out


def slice_patches_transpose(patches, shape, n, margin):
    """Adjoint of slice_patches: scatter-add windows, then drop the padding.

    This is what jax.linear_transpose(slice_patches) computes
    (jubik/convolve.py:138-143).  ``shape`` is the full shape of the
    un-sliced array.
    """
    patches = np.asarray(patches)
    dx, dy, hx, hy, wx, wy = patch_geometry(shape, n, margin)
    lead = tuple(shape[:-2])
    canvas = np.zeros(lead + (shape[-2] + 2 * hx, shape[-1] + 2 * hy),
                      dtype=patches.dtype)
    sx = window_starts(n, dx, canvas.shape[-2], wx)
    sy = window_starts(n, dy, canvas.shape[-1], wy)
    for p, (ix, iy) in enumerate(patch_order(n)):
        canvas[..., sx[ix]:sx[ix] + wx, sy[iy]:sy[iy] + wy] += patches[p]
    return canvas[..., hx:hx + shape[-2], hy:hy + shape[-1]]


def _fftn(a, axes):
    return sfft.fftn(a, axes=axes, workers=_WORKERS)


def _ifftn(a, axes):
    return sfft.ifftn(a, axes=axes, workers=_WORKERS)


def convolve(x, y, domain, axes, verbose=True):
    """FFT convolution, jubik/convolve.py:148-176.

    Returns dvol * Re(ifftn(fftn(x) * fftn(y))).  When y has more axes than
    x, x is broadcast over the axis just before the transformed ones and
    the reference prints a notice.
    """
    dvol = float(reduce(lambda a, b: a * b,
                        [domain.distances[i] for i in axes]))
    hx = _fftn(np.asarray(x), axes)
    hy = _fftn(np.asarray(y), axes)
    if np.ndim(y) > np.ndim(x):
        if verbose:
            print("kernel_shape:", np.shape(x))
            print("signal_shape:", np.shape(y))
            print("Dimension Inconsistency. Broadcasting PSFs")
        prod = hx[..., None, :, :] * hy
    else:
        prod = hx * hy
    return (dvol * _ifftn(prod, axes)).real


def integrate(x, domain, axes):
    """dvol * sum(x), jubik/convolve.py:179-184."""
    total = np.sum(x, axis=tuple(axes))
    dvol = float(reduce(lambda a, b: a * b,
                        [domain.distances[i] for i in axes]))
    return dvol * total


def linpatch_geometry(shape, n, margin):
    """dx, window w, FFT length L and kernel cut of linpatch_convolve."""
    dx = int(shape[-2] / n)
    dy = int(shape[-1] / n)
    cx = (shape[-2] - 2 * dx) // 2
    cy = (shape[-1] - 2 * dy) // 2
    return dict(dx=dx, dy=dy, cx=cx, cy=cy, wx=2 * dx, wy=2 * dy,
                Lx=2 * dx + 2 * margin, Ly=2 * dy + 2 * margin)


def cut_kernel(kernel, shape, n):
    """fftshift + centred cut of the full-frame kernel, convolve.py:104-111.

    A cut of zero (n = 2 with an even frame) gives an empty window exactly
    as the reference's ``[c:-c]`` slice does.
    """
    g = linpatch_geometry(shape, n, 0)
    rolled = np.fft.fftshift(np.asarray(kernel), axes=(-2, -1))
    cx, cy = g["cx"], g["cy"]
    return rolled[..., cx:-cx if cx else 0, cy:-cy if cy else 0]


def prepare_kernel_window(window, margin, dvol):
    """Pad, ifftshift and normalise a cut window, convolve.py:113-131."""
    window = np.asarray(window, dtype=np.float64)
    pad = [(0, 0)] * (window.ndim - 2) + [(margin, margin)] * 2
    padded = np.pad(window, pad)
    back = np.fft.ifftshift(padded, axes=(-2, -1))
    norm = back.sum(axis=(-2, -1)) * dvol
    return back / norm[..., None, None]


def _spectra(normed_kernel, padded_signal):
    hk = _fftn(normed_kernel, (-2, -1))
    if padded_signal.ndim > normed_kernel.ndim:
        hk = hk[..., None, :, :]
    return hk


def linpatch_convolve_window(x, domain, window, n, margin, verbose=False):
    """linpatch_convolve with the kernel given as its cut window.

    ``window`` has shape (P, [E,] 2dx, 2dy) and equals
    fftshift(kernel)[..., cx:-cx, cy:-cy] of the full-frame kernel.
    """
    from .data import Domain
    if not isinstance(domain, Domain):
        raise ValueError("domain has to be an instance of jubik.Domain")
    shape = tuple(domain.shape)
    x = np.asarray(x, dtype=np.float64)
    slices = slice_patches(x, shape, n, 0)
    weights = bilinear_weights(slices.shape[-2:])
    pad = [(0, 0)] * (slices.ndim - 2) + [(margin, margin)] * 2
    padded = np.pad(weights * slices, pad)
    dvol = domain.distances[-2] * domain.distances[-1]
    normed = prepare_kernel_window(window, margin, dvol)
    ndom = Domain((1,) + shape, (None,) + tuple(domain.distances))
    conv = convolve(normed, padded, ndom, axes=(-2, -1), verbose=verbose)
    padded_shape = tuple(shape[:-2]) + (shape[-2] + 2 * margin,
                                        shape[-1] + 2 * margin)
    canvas = slice_patches_transpose(conv, padded_shape, n, margin)
    return canvas[..., margin:-margin if margin else 0,
                  margin:-margin if margin else 0]


def linpatch_convolve(x, domain, kernel, n_patches_per_axis, margin,
                      verbose=False):
    """Spatially variant PSF by linear patching, jubik/convolve.py:66-145.

    ``kernel`` is the full-frame kernel array of shape (P, [E,] H, W).
    """
    from .data import Domain
    if not isinstance(domain, Domain):
        raise ValueError("domain has to be an instance of jubik.Domain")
    window = cut_kernel(kernel, domain.shape, n_patches_per_axis)
    return linpatch_convolve_window(x, domain, window, n_patches_per_axis,
                                    margin, verbose=verbose)


def linpatch_adjoint_window(g, domain, window, n, margin):
    """Hand-written adjoint of linpatch_convolve_window with respect to x.

    The chain is the transpose of each forward step, in reverse:
    embed the cotangent in the margin canvas, cut the L x L windows (the
    transpose of overlap-add), correlate with the kernel (conjugate
    spectrum), crop the margins, apply the tent weights and overlap-add the
    windows back onto the image (the transpose of slicing).
    """
    shape = tuple(domain.shape)
    g = np.asarray(g, dtype=np.float64)
    pad = [(0, 0)] * (g.ndim - 2) + [(margin, margin)] * 2
    gpad = np.pad(g, pad)
    padded_shape = tuple(shape[:-2]) + (shape[-2] + 2 * margin,
                                        shape[-1] + 2 * margin)
    gwin = slice_patches(gpad, padded_shape, n, margin)
    dvol = domain.distances[-2] * domain.distances[-1]
    normed = prepare_kernel_window(window, margin, dvol)
    hk = _fftn(normed, (-2, -1))
    hg = _fftn(gwin, (-2, -1))
    if gwin.ndim > normed.ndim:
        hk = hk[..., None, :, :]
        hg_sum = hg * np.conj(hk)
    else:
        hg_sum = hg * np.conj(hk)
    zbar = (dvol * _ifftn(hg_sum, (-2, -1))).real
    if margin:
        zbar = zbar[..., margin:-margin, margin:-margin]
    weights = bilinear_weights(zbar.shape[-2:])
    return slice_patches_transpose(weights * zbar, shape, n, 0)


def linpatch_adjoint(g, domain, kernel, n_patches_per_axis, margin):
    """Adjoint of linpatch_convolve for a full-frame kernel."""
    window = cut_kernel(kernel, domain.shape, n_patches_per_axis)
    return linpatch_adjoint_window(g, domain, window, n_patches_per_axis,
                                   margin)


def window_to_full_frame(window, shape):
    """Embed a cut window (P, [E,] w, w) in a zero full frame (.., H, W).

    Inverse of cut_kernel: the window's centre (index w/2) lands on the
    frame's origin, as the reference's fftshift/cut pair expects.
    """
    window = np.asarray(window)
    H, W = shape[-2], shape[-1]
    wx, wy = window.shape[-2], window.shape[-1]
    full = np.zeros(window.shape[:-2] + (H, W), dtype=window.dtype)
    cx, cy = (H - wx) // 2, (W - wy) // 2
    full[..., cx:cx + wx, cy:cy + wy] = window
    return np.fft.ifftshift(full, axes=(-2, -1))
