"""Batched 128-point DFTs two ways (the FP32-FFT vs split-BF16 matrix-core
comparison of the 2048^2 x 4 energy configuration): ``method="fft"`` is the
FP32 Stockham FFT the PSF kernels use, ``method="mfma"`` computes Y = F X as
real GEMMs on the matrix cores with each product split into BF16 halves
(hi*hi + hi*lo + lo*hi, FP32 accumulation)."""
import ctypes

import torch

from . import _device as dev
from . import _lib


class DFT128:
    def __init__(self):
        dev.require_device()
        self._plan = ctypes.c_void_p()
        _lib.call("jubik_dft128_create", ctypes.byref(self._plan))

    def __call__(self, x, method="fft", out=None):
        """Forward DFT along the last axis of a (B, 128) complex64 device
        tensor, B a multiple of 256."""
        if method not in ("fft", "mfma"):
            raise ValueError("method must be 'fft' or 'mfma'")
        xt = torch.as_tensor(x, device="cuda").to(torch.complex64).contiguous()
        if xt.dim() != 2 or xt.shape[1] != 128:
            raise ValueError("x must have shape (B, 128)")
        y = torch.empty_like(xt) if out is None else out
        _lib.call("jubik_dft128_run", self._plan, 0 if method == "fft" else 1, dev.ptr(xt),
                  dev.ptr(y), xt.shape[0], dev.stream())
        return y

    def __del__(self):
        p = getattr(self, "_plan", None)
        if p is not None and p.value:
            try:
                _lib.load().jubik_dft128_destroy(p)
            except Exception:
                pass
            self._plan = None
