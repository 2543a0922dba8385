"""Plain FP64 GPU path: the reference's op order with torch ops and autograd.

This is the comparison baseline that bench.py times beside the fused
kernels; it is not part of the library.  It restates jubik/convolve.py:28-176
(pad, window slicing, tent weights, margin pad, per-call kernel fftshift/
cut/pad/ifftshift/normalise, complex128 FFT convolution, overlap-add, crop),
the exposure/read-out of jubik/response.py and the Poisson energy as
separate torch operations (hipFFT Z2Z for the transforms), and takes the
gradient by autograd, as JAX does for the reference.  Every intermediate is a
complex128 (P, E, L, L) array in HBM.

The window slicing is one gather over all patches (what XLA makes of the
vmapped ``lax.dynamic_slice`` at jubik/convolve.py:55-63) and the
overlap-add one scatter-add (the ``linear_transpose`` of that slicing at
jubik/convolve.py:138-143); autograd transposes each into the other.
"""
import numpy as np
import torch


def _tent(w, device):
    half = w // 2
    ramp = torch.as_tensor(np.linspace(0.0, 1.0, half), dtype=torch.float64, device=device)
    b = torch.cat([ramp, ramp.flip(0)])
    return b[:, None] * b[None, :]


def _starts(n, d, padded, win):
    return [min(i * d, padded - win) for i in range(n)]


class PlainFP64Model:
    """NLL(xi) = sum(lam - d log lam), lam = expo * linpatch(a * s(xi), K)."""

    def __init__(self, amp_ref, amp_alpha, zm, mu, dnu, kernel_full, n, margin, pixel_area,
                 exposure, keep, data, prepare_each_call=True, device="cuda"):
        dev = device
        f64 = torch.float64
        self.amp_ref = torch.as_tensor(amp_ref, dtype=f64, device=dev)
        self.amp_alpha = None if amp_alpha is None else torch.as_tensor(amp_alpha, dtype=f64,
                                                                         device=dev)
        self.zm, self.mu = float(zm), float(mu)
        self.dnu = torch.as_tensor(np.asarray(dnu, dtype=np.float64), device=dev)[:, None, None]
        self.kernel = torch.as_tensor(kernel_full, dtype=f64, device=dev)  # (P, E, H, W)
        self.n, self.m = int(n), int(margin)
        self.a = float(pixel_area)
        self.expo = torch.as_tensor(exposure, dtype=f64, device=dev)
        self.keep = torch.as_tensor(keep, dtype=torch.bool, device=dev)
        self.data = torch.as_tensor(data, dtype=f64, device=dev)
        self.prepare_each_call = prepare_each_call
        E, H, W = self.expo.shape
        self.shape = (E, H, W)
        self.dx = H // self.n
        self.w = 2 * self.dx
        self.h = self.dx // 2
        self.L = self.w + 2 * self.m
        self.weights = _tent(self.w, dev)
        n, w, h, m, L = self.n, self.w, self.h, self.m, self.L
        # Patch p = iy * n + ix starts at (sx[ix], sy[iy]) in the padded image
        # (convolve.py:59-63); one index tensor per axis serves the gather,
        # one flat canvas index the scatter.
        sx = torch.as_tensor(_starts(n, self.dx, H + 2 * h, w), device=dev)
        sy = torch.as_tensor(_starts(n, self.dx, W + 2 * h, w), device=dev)
        p = torch.arange(n * n, device=dev)
        px, py = sx[p % n], sy[p // n]
        self._rows = (px[:, None] + torch.arange(w, device=dev)[None, :])[:, :, None]
        self._cols = (py[:, None] + torch.arange(w, device=dev)[None, :])[:, None, :]
        Wc = W + 2 * m + 2 * h
        self._canvas_shape = (E, H + 2 * m + 2 * h, Wc)
        ar = torch.arange(L, device=dev)
        self._flat = ((px[:, None, None] + ar[None, :, None]) * Wc +
                      (py[:, None, None] + ar[None, None, :])).reshape(-1)
        self._khat = None if prepare_each_call else self._kernel_spectrum()

    def _kernel_spectrum(self):
        H, W = self.shape[-2:]
        c = (H - self.w) // 2
        rolled = torch.fft.fftshift(self.kernel, dim=(-2, -1))
        cut = rolled[..., c:H - c, c:W - c]
        padded = torch.nn.functional.pad(cut, (self.m,) * 4)
        back = torch.fft.ifftshift(padded, dim=(-2, -1))
        normed = back / back.sum(dim=(-2, -1), keepdim=True)
        return torch.fft.fftn(normed, dim=(-2, -1))

    def sky(self, xi_ref, xi_alpha):
        f = torch.fft.fftn(self.amp_ref * xi_ref)
        log_s = self.zm + (f.real + f.imag)[None]
        if self.amp_alpha is not None:
            g = torch.fft.fftn(self.amp_alpha * xi_alpha)
            log_s = log_s + ((g.real + g.imag)[None] + self.mu) * self.dnu
        E, H, W = self.shape
        return torch.exp(log_s)[:, :H, :W]

    def linpatch(self, x):
        E, H, W = self.shape
        h, m = self.h, self.m
        xp = torch.nn.functional.pad(x, (h, h, h, h))
        slices = xp[:, self._rows, self._cols].transpose(0, 1)  # (P, E, w, w), one gather
        z = torch.nn.functional.pad(self.weights * slices, (m, m, m, m))
        khat = self._kernel_spectrum() if self.prepare_each_call else self._khat
        y = torch.fft.ifftn(khat * torch.fft.fftn(z, dim=(-2, -1)), dim=(-2, -1)).real
        vals = y.transpose(0, 1).reshape(E, -1)  # (E, P * L * L)
        canvas = torch.zeros((E, self._canvas_shape[1] * self._canvas_shape[2]), dtype=y.dtype,
                             device=y.device).index_add(1, self._flat, vals)  # one scatter-add
        canvas = canvas.reshape(self._canvas_shape)
        return canvas[:, h + m:h + m + H, h + m:h + m + W]

    def nll(self, xi_ref, xi_alpha=None):
        s = self.sky(xi_ref, xi_alpha)
        lam = (self.expo * self.linpatch(self.a * s))[self.keep]
        return lam.sum() - (self.data * torch.log(lam)).sum()

    def nll_and_grad(self, xi_ref, xi_alpha=None):
        xr = xi_ref.detach().to(torch.float64).requires_grad_(True)
        xa = None
        inputs = [xr]
        if self.amp_alpha is not None:
            xa = xi_alpha.detach().to(torch.float64).requires_grad_(True)
            inputs.append(xa)
        val = self.nll(xr, xa)
        grads = torch.autograd.grad(val, inputs)
        return val.detach(), grads[0], (grads[1] if len(grads) > 1 else None)
