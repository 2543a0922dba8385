"""Correlated-field GP sky on the GPU with a fixed amplitude field.

Implements the single-spectral-behaviour sky of
jubik/sky_model/multifrequency/spectral_product_mf_sky.py:171-219 (and the
one-frequency reference distribution, 281-298; the diffuse exp(CF) of
jubik/sky_models.py:306-341 is the same arithmetic):

    s_e = exp(zm + HT(amp_ref xi_ref) + (HT(amp_alpha xi_alpha) + mu) dnu_e)

with amp = hdvol * fluctuations * A[pd], A[0] = 0 and HT nifty.re's hartley.
``powerlaw_amplitude`` builds a closed-form amplitude field for synthetic
workloads (the reference infers A; here it is an input).
"""
import ctypes

import numpy as np
import torch

from . import _device as dev
from . import _lib


def powerlaw_amplitude(N, distance=1.0, slope=-4.0, fluctuations=1.0):
    """amp = hdvol * f * A[pd] for a power spectrum P(k) = |k|^slope.

    Normalised so the pixel variance of HT(amp * xi), xi ~ N(0, 1), is
    fluctuations**2 (averaged over pixels): sum_k amp_k^2 = f^2.  A[0] = 0.
    """
    k = np.fft.fftfreq(N, d=distance)
    kk = np.sqrt(k[:, None] ** 2 + k[None, :] ** 2)
    amp = np.zeros((N, N))
    nz = kk > 0
    amp[nz] = kk[nz] ** (slope / 2.0)
    amp *= fluctuations / np.sqrt(np.sum(amp ** 2))
    return amp


def powerlaw_amplitude_grad(N, g_amp, distance=1.0, slope=-4.0, fluctuations=1.0):
    """Gradients of a scalar with respect to the power-law amplitude's
    hyperparameters (slope, fluctuations), given its gradient g_amp with
    respect to the amplitude field ``powerlaw_amplitude`` builds (for
    instance g_amp_ref of ``XrayModel.nll_and_grad_hyper``).

    With amp = f u / |u|, u_k = |k|^(slope / 2):
    d amp_k / d f = amp_k / f and
    d amp_k / d slope = amp_k (ln|k_k| / 2 - sum_j amp_j^2 ln|k_j| / (2 f^2)).
    Returns (g_slope, g_fluctuations) as floats (float64 sums on the host).
    """
    g = g_amp.detach().double().cpu().numpy() if isinstance(g_amp, torch.Tensor) else \
        np.asarray(g_amp, dtype=np.float64)
    if g.shape != (N, N):
        raise ValueError(f"g_amp has shape {g.shape}, expected {(N, N)}")
    amp = powerlaw_amplitude(N, distance, slope, fluctuations)
    k = np.fft.fftfreq(N, d=distance)
    kk = np.sqrt(k[:, None] ** 2 + k[None, :] ** 2)
    half_log = np.zeros((N, N))
    nz = kk > 0
    half_log[nz] = 0.5 * np.log(kk[nz])
    c = float(np.sum(amp ** 2 * half_log)) / fluctuations ** 2
    ga = g * amp
    return float(np.sum(ga * half_log) - c * np.sum(ga)), float(np.sum(ga)) / fluctuations


class CorrelatedFieldSky:
    """GP sky plan: N x N GP grid cropped at the origin to (H, W)."""

    def __init__(self, N, shape, amp_ref, amp_alpha=None, zm=0.0, mu=0.0, dnu=None,
                 dtype=torch.float32):
        dev.require_device()
        self.N = int(N)
        self.H, self.W = int(shape[-2]), int(shape[-1])
        self.has_alpha = amp_alpha is not None
        self.dnu = np.zeros(1) if dnu is None else np.asarray(dnu, dtype=np.float64).ravel()
        self.E = len(self.dnu) if self.has_alpha else 1
        self.dtype = dtype
        self.zm, self.mu = float(zm), float(mu)
        self.amp_ref = dev.to_device(amp_ref, dtype)
        self.amp_alpha = dev.to_device(amp_alpha, dtype) if self.has_alpha else None
        self.dnu_dev = dev.to_device(self.dnu, dtype)
        for a in (self.amp_ref, self.amp_alpha):
            if a is not None and tuple(a.shape) != (self.N, self.N):
                raise ValueError(f"amplitude fields must be ({self.N}, {self.N})")
        self.version = 0          # bumped by every set_parameters call
        self.offsets_version = 0  # bumped when zm or mu change
        self._plan = ctypes.c_void_p()
        _lib.call("jubik_gp_create", ctypes.byref(self._plan), dev.code(dtype), self.N, self.H,
                  self.W, self.E, int(self.has_alpha))

    def __del__(self):
        plan = getattr(self, "_plan", None)
        if plan is not None and plan.value:
            try:
                _lib.load().jubik_gp_destroy(plan)
            except Exception:
                pass
            self._plan = None

    def set_parameters(self, amp_ref=None, amp_alpha=None, zm=None, mu=None):
        """Replace the amplitude fields and/or offsets (the outputs of the
        reference's amplitude models and zero-mode priors, which change as
        their hyperparameters are inferred); shapes and dtype are fixed.

        The amplitude fields are overwritten in place, so their device
        addresses stay valid for captured graphs, which read them at every
        launch.  The offsets are launch arguments: a graph captured before
        they change refuses to launch (``offsets_version``).  Every call
        bumps ``version``; a model's metric base point taken at an older
        version is refused."""
        new = []
        for name, a in (("amp_ref", amp_ref), ("amp_alpha", amp_alpha)):
            if a is None:
                continue
            if name == "amp_alpha" and not self.has_alpha:
                raise ValueError("this sky has no spectral-index field")
            t = dev.to_device(a, self.dtype)
            if tuple(t.shape) != (self.N, self.N):
                raise ValueError(f"amplitude fields must be ({self.N}, {self.N})")
            new.append((getattr(self, name), t))
        for dst, t in new:
            dst.copy_(t)
        if (zm is not None and float(zm) != self.zm) or (mu is not None and float(mu) != self.mu):
            self.offsets_version += 1
        if zm is not None:
            self.zm = float(zm)
        if mu is not None:
            self.mu = float(mu)
        self.version += 1

    @property
    def shape(self):
        return (self.E, self.H, self.W)

    def __call__(self, xi_ref, xi_alpha=None):
        """s of shape (E, H, W)."""
        xr = dev.to_device(xi_ref, self.dtype)
        xa = dev.to_device(xi_alpha, self.dtype) if self.has_alpha else None
        if self.has_alpha and xi_alpha is None:
            raise ValueError("xi_alpha is required for a spectral sky")
        s = torch.empty(self.shape, dtype=self.dtype, device="cuda")
        _lib.call("jubik_gp_forward", self._plan, dev.ptr(self.amp_ref), dev.ptr(self.amp_alpha),
                  dev.ptr(xr), dev.ptr(xa), self.zm, self.mu, dev.ptr(self.dnu_dev), dev.ptr(s),
                  dev.stream())
        return dev.like_input(s, xi_ref)

    def vjp(self, sky, cot):
        """Cotangents of (xi_ref, xi_alpha) for the cotangent ``cot`` of s."""
        st = dev.to_device(sky, self.dtype)
        ct = dev.to_device(cot, self.dtype)
        if tuple(st.shape) != self.shape or tuple(ct.shape) != self.shape:
            raise ValueError(f"sky and cotangent must have shape {self.shape}")
        g_ref = torch.empty((self.N, self.N), dtype=self.dtype, device="cuda")
        g_alpha = torch.empty_like(g_ref) if self.has_alpha else None
        _lib.call("jubik_gp_vjp", self._plan, dev.ptr(self.amp_ref), dev.ptr(self.amp_alpha),
                  dev.ptr(self.dnu_dev), dev.ptr(st), dev.ptr(ct), dev.ptr(g_ref),
                  dev.ptr(g_alpha), dev.stream())
        g_ref = dev.like_input(g_ref, cot)
        return (g_ref, dev.like_input(g_alpha, cot)) if self.has_alpha else (g_ref, None)
