"""Exposure and read-out: drop-in for jubik/response.py on the GPU.

The builders keep the reference's host-side behaviour exactly (argument
checks, in-place cut of the caller's array, key handling); the closures
they return run as HIP kernels (broadcast multiply, gather by the C-order
indices of the kept pixels).
"""
import numpy as np
import torch

from . import _device as dev
from . import _lib
from .data import Vector


def build_exposure_function(exposures, exposure_cut=None):
    """jubik/response.py:12-45: returns ``x -> exposures * x``.

    Entries below ``exposure_cut`` are zeroed in place in ``exposures`` (the
    reference mutates the caller's array, which Chandra relies on).  Like
    the reference's closure, each call multiplies by the array as it is at
    that call: later in-place edits of ``exposures`` are seen (a device
    tensor of the working dtype is used without a copy; a host array is
    copied to the device per call).
    """
    if exposure_cut is not None:
        if exposure_cut < 0:
            raise ValueError("exposure_cut should be non-negative or None!")
        exposures[exposures < exposure_cut] = 0

    def exposure(x):
        dtype = dev.float_dtype(exposures, x)
        dev.require_device()
        e = dev.to_device(exposures, dtype)
        xt = dev.to_device(x, dtype)
        try:
            oshape = tuple(torch.broadcast_shapes(tuple(e.shape), tuple(xt.shape)))
        except RuntimeError as err:
            raise ValueError(str(err)) from None
        ev, xv = e, xt
        # The kernel broadcasts operands whose shape is a suffix of the
        # output shape; any other broadcast is materialised first.
        if tuple(ev.shape) != oshape[len(oshape) - ev.dim():]:
            ev = ev.expand(oshape).contiguous()
        if tuple(xv.shape) != oshape[len(oshape) - xv.dim():]:
            xv = xv.expand(oshape).contiguous()
        out = torch.empty(oshape, dtype=dtype, device="cuda")
        _lib.call("jubik_exposure_apply", dev.code(dtype), dev.ptr(ev), max(ev.numel(), 1),
                  dev.ptr(xv), max(xv.numel(), 1), dev.ptr(out), out.numel(), dev.stream())
        return dev.like_input(out, x)

    return exposure


def build_readout_function(flags, threshold=None, keys=None):
    """jubik/response.py:48-112: returns ``x -> Vector({key_i: x[i][~mask[i]]})``.

    ``mask = (flags == 0)`` after zeroing the flags below ``threshold`` in
    place.  The read-out is a gather by the precomputed C-order indices of
    the kept pixels of every module.
    """
    if threshold is not None and threshold < 0:
        raise ValueError("threshold should be positive!")
    if threshold is not None:
        flags[flags < threshold] = 0
    mask = np.asarray(flags) == 0

    if keys is None:
        keys = ['masked input']
    elif len(keys) != flags.shape[0]:
        raise ValueError("length of keys should match the number of flag maps.")
    index_cache = {}

    def _indices(i):
        dkey = (i, torch.cuda.current_device())
        if dkey not in index_cache:
            flat = np.flatnonzero(~mask[i].reshape(-1)).astype(np.int64)
            index_cache[dkey] = torch.as_tensor(flat, device="cuda")
        return index_cache[dkey]

    def _apply_readout(x):
        if len(mask.shape) != 4:
            raise ValueError("flags should have shape (n, m, q, l)!")
        if len(x.shape) != 4:
            raise ValueError("input should have shape (n, m, q, l)!")
        dev.require_device()
        dtype = dev.float_dtype(x)
        xt = dev.to_device(x, dtype)
        per = int(np.prod(xt.shape[1:]))
        out = {}
        for i, key in enumerate(keys):
            idx = _indices(i)
            v = torch.empty(idx.numel(), dtype=dtype, device="cuda")
            src = xt[i].reshape(-1)
            if src.numel() != per:
                raise ValueError("input module shape does not match the flags")
            _lib.call("jubik_readout_gather", dev.code(dtype), dev.ptr(src), dev.ptr(idx),
                      idx.numel(), dev.ptr(v), dev.stream())
            out[key] = dev.like_input(v, x)
        return Vector(out)

    def _transpose(vec, like_shape=None):
        """Scatter read-out vectors back into a zero (M, E, H, W) array."""
        dev.require_device()
        shape = mask.shape
        first = vec[keys[0]]
        dtype = dev.float_dtype(first)
        out = torch.empty(shape, dtype=dtype, device="cuda")
        per = int(np.prod(shape[1:]))
        for i, key in enumerate(keys):
            idx = _indices(i)
            v = dev.to_device(vec[key], dtype)
            dst = out[i].reshape(-1)
            _lib.call("jubik_readout_scatter", dev.code(dtype), dev.ptr(v), dev.ptr(idx),
                      idx.numel(), per, dev.ptr(dst), dev.stream())
        return dev.like_input(out, first)

    _apply_readout.transpose = _transpose
    _apply_readout.mask = mask
    _apply_readout.keys = tuple(keys)
    return _apply_readout
