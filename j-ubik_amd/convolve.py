"""Spatially variant PSF: drop-in for jubik/convolve.py on the GPU.

``linpatch_convolve(x, domain, kernel, n_patches_per_axis, margin)`` keeps
the reference signature (jubik/convolve.py:66) and error behaviour; the
work runs in libjubik_hip.so (row FFT pass, column FFT pass with the kernel
spectrum, gather-form overlap-add).  ``LinpatchOperator`` is the reusable
plan (device workspace + prepared kernel spectra) behind it, with the
adjoint used for gradients.
"""
import ctypes
import json
import math
from pathlib import Path

import numpy as np
import torch

from . import _device as dev
from . import _lib
from .data import is_domain


def _geometry(shape, n, margin):
    H, W = int(shape[-2]), int(shape[-1])
    dx, dy = int(H / n), int(W / n)
    return dict(H=H, W=W, dx=dx, dy=dy, w=2 * dx, L=2 * dx + 2 * margin)


class LinpatchOperator:
    """Plan for linpatch_convolve on one geometry.

    shape        sky shape (*lead, H, W); lead axes (energies) are batched.
    n_modules    number of kernel sets applied to the same input (the vmap
                 over modules of jubik/instruments/*/..._response.py).
    kernel_per_energy  kernels carry the lead axes (else one kernel per
                 patch serves every energy, convolve.py:167-171).
    x_shared     one input for all modules (else one input per module).
    dtype        torch.float32 or torch.float64.
    """

    def __init__(self, shape, n_patches_per_axis, margin, n_modules=1,
                 kernel_per_energy=True, x_shared=True, dtype=torch.float32):
        dev.require_device()
        self.shape = tuple(int(s) for s in shape)
        self.lead = self.shape[:-2]
        self.E = int(np.prod(self.lead)) if self.lead else 1
        self.H, self.W = self.shape[-2:]
        self.n = int(n_patches_per_axis)
        self.margin = int(margin)
        self.M = int(n_modules)
        self.kernel_per_energy = bool(kernel_per_energy)
        self.x_shared = bool(x_shared)
        self.dtype = dtype
        self._plan = ctypes.c_void_p()
        _lib.call("jubik_linpatch_create", ctypes.byref(self._plan), dev.code(dtype),
                  self.H, self.W, self.E, self.M, self.n, self.margin,
                  int(self.kernel_per_energy), int(self.x_shared))
        info = (ctypes.c_int64 * 8)()
        _lib.call("jubik_linpatch_info", self._plan, info)
        self.dx, self.window, self.L, self.P, self.pad, self.J = (int(v) for v in info[:6])
        self.workspace_bytes, self.spectra_bytes = int(info[6]), int(info[7])
        self._kernel_ref = None
        self.col_div, self.row_div = 1, 1
        self.transform = "fft"

    def __del__(self):
        plan = getattr(self, "_plan", None)
        if plan is not None and plan.value:
            try:
                _lib.load().jubik_linpatch_destroy(plan)
            except Exception:
                pass
            self._plan = None

    # -------------------------------------------------------------- kernels
    def kernel_shape(self, layout="full"):
        side = (self.H, self.W) if layout == "full" else (self.window, self.window)
        ek = self.lead if self.kernel_per_energy else ()
        mods = (self.M,) if self.M > 1 else ()
        return mods + (self.P,) + ek + side

    def set_kernel(self, kernel, layout="full"):
        """Prepare the kernel spectra (convolve.py:104-131).

        layout "full": (M, P, [lead,] H, W) full-frame kernels as the
        reference takes them; "window": the cut (M, P, [lead,] 2dx, 2dx)
        windows, i.e. fftshift(kernel)[..., c:-c, c:-c].
        """
        lay = {"full": _lib.KERNEL_FULL, "window": _lib.KERNEL_WINDOW}[layout]
        k = dev.to_device(kernel, self.dtype)
        expect = self.kernel_shape(layout)
        if tuple(k.shape) != expect and not (self.M == 1 and tuple(k.shape) == expect[1:]):
            raise ValueError(f"kernel has shape {tuple(k.shape)}, expected {expect}")
        _lib.call("jubik_linpatch_set_kernel", self._plan, dev.ptr(k), lay, dev.stream())
        self._kernel_ref = k  # keep alive until the preparation has run
        return self

    # -------------------------------------------------------------- apply
    def forward(self, x, scale=1.0, out=None):
        """PSF(scale * x): (M, *lead, H, W) (module axis dropped when M = 1)."""
        xt = dev.to_device(x, self.dtype)
        want = self.shape if self.x_shared else (self.M,) + self.shape
        if tuple(xt.shape) != want:
            raise ValueError(f"input has shape {tuple(xt.shape)}, expected {want}")
        oshape = ((self.M,) if self.M > 1 else ()) + self.shape
        y = out if out is not None else torch.empty(oshape, dtype=self.dtype, device="cuda")
        _lib.call("jubik_linpatch_forward", self._plan, dev.ptr(xt), dev.ptr(y),
                  float(scale), dev.stream())
        return y

    def adjoint(self, gy, scale=1.0, out=None):
        """scale * PSF^T(gy), summed over modules when the input is shared."""
        gt = dev.to_device(gy, self.dtype)
        ishape = ((self.M,) if self.M > 1 else ()) + self.shape
        if tuple(gt.shape) != ishape:
            raise ValueError(f"cotangent has shape {tuple(gt.shape)}, expected {ishape}")
        oshape = self.shape if self.x_shared else (self.M,) + self.shape
        gx = out if out is not None else torch.empty(oshape, dtype=self.dtype, device="cuda")
        _lib.call("jubik_linpatch_adjoint", self._plan, dev.ptr(gt), dev.ptr(gx),
                  float(scale), dev.stream())
        return gx

    def __call__(self, x):
        return self.forward(x)

    # Linearisation protocol of the likelihood composition (Poissonian.amend):
    # the operator is linear, so its JVP is the forward and its VJP the adjoint.
    def jvp(self, x, t):
        return self.forward(t)

    def vjp(self, x, ct):
        return self.adjoint(ct)

    # -------------------------------------------------------------- tuning
    def tune_key(self, samples=1):
        """Key of this geometry in a tile cache: device architecture, dtype,
        FFT length, grid, energies, modules, patches, samples per pass."""
        arch = getattr(torch.cuda.get_device_properties(torch.cuda.current_device()),
                       "gcnArchName", "gpu")
        dt = "f32" if self.dtype == torch.float32 else "f64"
        return (f"{arch}/{dt}/L{self.L}/{self.H}x{self.W}/E{self.E}/M{self.M}/P{self.P}"
                f"/S{int(samples)}/xs{int(self.x_shared)}/ke{int(self.kernel_per_energy)}")

    def set_tile(self, col_div, row_div=1):
        """Column- and row-pass tile divisors: 1, 2 or 4 each (results agree
        to rounding)."""
        _lib.call("jubik_linpatch_set_tile", self._plan, int(col_div), int(row_div))
        self.col_div, self.row_div = int(col_div), int(row_div)
        return self

    TRANSFORMS = ("fft", "mfma_bf16x3")

    def set_transform(self, kind):
        """Transform of the column passes: "fft" (the plan's FFT, default) or
        "mfma_bf16x3" (float32 plans at FFT length 384: each column DFT as
        three 128-point DFTs on the matrix cores with BF16 hi/lo split
        products, hi*hi + hi*lo + lo*hi, FP32 accumulation, and an FP32
        radix-3 step).  Applies to the models built on this operator too."""
        if kind not in self.TRANSFORMS:
            raise ValueError(f"transform must be one of {self.TRANSFORMS}")
        _lib.call("jubik_linpatch_set_transform", self._plan, self.TRANSFORMS.index(kind))
        self.transform = kind
        return self

    def autotune(self, reps=20, cache=None):
        """Pick the fastest column- and row-pass tiles for this geometry by
        timing one forward + adjoint per candidate (the kernel must be set).
        ``cache``: a JSON file of earlier choices, read first and updated."""
        key = self.tune_key()
        hit = tile_cache_get(cache, key)
        if hit is not None:
            self.set_tile(*hit)
            return {"col_div": hit[0], "row_div": hit[1], "ms": None, "cached": True}
        col, row = ctypes.c_int(), ctypes.c_int()
        ms = (ctypes.c_double * 6)()
        _lib.call("jubik_linpatch_autotune", self._plan, int(reps), ctypes.byref(col),
                  ctypes.byref(row), ms, dev.stream())
        self.col_div, self.row_div = int(col.value), int(row.value)
        tile_cache_put(cache, key, (self.col_div, self.row_div))
        return tune_result(self.col_div, self.row_div, ms)


def tune_result(col_div, row_div, ms):
    """The tuner's pick and its timings (ms per forward + adjoint) by divisor."""
    return {"col_div": col_div, "row_div": row_div, "cached": False,
            "ms": {"col": {d: t for d, t in zip((1, 2, 4), ms[:3]) if t >= 0},
                   "row": {d: t for d, t in zip((1, 2, 4), ms[3:6]) if t >= 0}}}


def tile_cache_get(cache, key):
    """(col_div, row_div) stored under ``key`` in the JSON file ``cache``,
    or None."""
    if cache is None:
        return None
    try:
        table = json.loads(Path(cache).read_text())
    except FileNotFoundError:
        return None
    hit = table.get(key)
    return None if hit is None else (int(hit[0]), int(hit[1]))


def tile_cache_put(cache, key, tiles):
    """Store (col_div, row_div) under ``key`` in the JSON file ``cache``."""
    if cache is None:
        return
    path = Path(cache)
    try:
        table = json.loads(path.read_text())
    except FileNotFoundError:
        table = {}
    table[key] = [int(tiles[0]), int(tiles[1])]
    tmp = path.with_suffix(path.suffix + ".tmp")
    tmp.write_text(json.dumps(table, indent=1, sort_keys=True) + "\n")
    tmp.replace(path)


_PLANS = {}


def _plan_for(shape, n, margin, M, kernel_per_energy, dtype):
    key = (tuple(shape), n, margin, M, kernel_per_energy, dtype, torch.cuda.current_device())
    op = _PLANS.get(key)
    if op is None:
        op = LinpatchOperator(shape, n, margin, n_modules=M,
                              kernel_per_energy=kernel_per_energy, dtype=dtype)
        _PLANS[key] = op
    return op


def _check_linpatch_args(x, domain, kernel, n, margin, modules):
    if not is_domain(domain):
        raise ValueError("domain has to be an instance of jubik.Domain")
    shape = tuple(int(s) for s in domain.shape)
    if tuple(x.shape[-2:]) != shape[-2:]:
        raise ValueError(f"input of shape {tuple(x.shape)} does not match domain {shape}")
    lead = tuple(x.shape[:-2])
    kshape = tuple(kernel.shape)
    if modules:
        kshape = kshape[1:]
    P = n * n
    if not kshape or kshape[0] != P:
        raise ValueError(f"kernel needs {P} patches on its first axis, got {tuple(kernel.shape)}")
    if kshape[-2:] != shape[-2:]:
        raise ValueError("kernel frames must have the shape of the domain")
    klead = kshape[1:-2]
    if klead == lead:
        per_energy = True
    elif klead == ():
        per_energy = False
    else:
        raise ValueError(f"kernel shape {tuple(kernel.shape)} does not broadcast to input "
                         f"{tuple(x.shape)}")
    return lead, per_energy


def _linpatch(x, domain, kernel, n, margin, modules):
    n = int(n)
    margin = int(margin)
    lead, per_energy = _check_linpatch_args(x, domain, kernel, n, margin, modules)
    M = int(kernel.shape[0]) if modules else 1
    dtype = dev.float_dtype(x, kernel)
    out_shape = ((M,) if modules else ()) + lead
    if margin == 0:
        # The reference crops [..., 0:-0, 0:-0]: an empty result.
        empty = torch.empty(out_shape + (0, 0), dtype=dtype)
        return empty if isinstance(x, torch.Tensor) else empty.numpy()
    if not per_energy and lead:
        g = _geometry(domain.shape, n, margin)
        P = n * n
        print("kernel_shape:", (P, g["L"], g["L"]))
        print("signal_shape:", (P,) + lead + (g["L"], g["L"]))
        print("Dimension Inconsistency. Broadcasting PSFs")
    dev.require_device()
    op = _plan_for(tuple(lead) + tuple(x.shape[-2:]), n, margin, M, per_energy, dtype)
    op.set_kernel(kernel, "full")
    y = op.forward(x)
    if modules and M == 1:
        y = y.unsqueeze(0)
    return dev.like_input(y, x)


def linpatch_convolve(x, domain, kernel, n_patches_per_axis, margin):
    """Functional linear-patching convolution (jubik/convolve.py:66-145).

    x (*lead, H, W); kernel (P, [*lead,] H, W) full-frame PSFs, one per
    patch, P = n_patches_per_axis ** 2.  float32 inputs run in FP32,
    anything else in FP64.
    """
    return _linpatch(x, domain, kernel, n_patches_per_axis, margin, modules=False)


def linpatch_convolve_modules(x, domain, kernels, n_patches_per_axis, margin):
    """``vmap(linpatch_convolve, in_axes=(None, None, 0, None, None))`` of the
    response builders (chandra_response.py:127-130, erosita_response.py:209-212):
    kernels (M, P, [*lead,] H, W) -> (M, *lead, H, W)."""
    return _linpatch(x, domain, kernels, n_patches_per_axis, margin, modules=True)


def slice_patches(x, shape, n_patches_per_axis, additional_margin):
    """Overlapping windows of x, jubik/convolve.py:28-63, on the GPU.

    Output (n^2, *x.shape[:-2], 2dx + 2m, 2dy + 2m) with dx = (H - 2m) // n;
    patch p = iy * n + ix starts at (ix * dx, iy * dy) of x zero-padded by
    dx // 2 (starts clamped as jax.lax.dynamic_slice does).
    """
    n, m = int(n_patches_per_axis), int(additional_margin)
    H, W = int(shape[-2]), int(shape[-1])
    if tuple(x.shape[-2:]) != (H, W):
        raise ValueError(f"input of shape {tuple(x.shape)} does not match shape {tuple(shape)}")
    dev.require_device()
    dtype = dev.float_dtype(x)
    xt = dev.to_device(x, dtype)
    lead = tuple(xt.shape[:-2])
    B = int(np.prod(lead)) if lead else 1
    dx, dy = int((H - 2 * m) / n), int((W - 2 * m) / n)
    out = torch.empty((n * n,) + lead + (2 * dx + 2 * m, 2 * dy + 2 * m), dtype=dtype,
                      device="cuda")
    _lib.call("jubik_slice_patches", dev.code(dtype), dev.ptr(xt), B, H, W, n, m, dev.ptr(out),
              dev.stream())
    return dev.like_input(out, x)


def slice_patches_transpose(patches, shape, n_patches_per_axis, additional_margin):
    """Transpose of slice_patches (what jax.linear_transpose gives,
    jubik/convolve.py:138-143): sums the windows back onto (*lead, H, W)."""
    n, m = int(n_patches_per_axis), int(additional_margin)
    H, W = int(shape[-2]), int(shape[-1])
    dev.require_device()
    dtype = dev.float_dtype(patches)
    pt = dev.to_device(patches, dtype)
    lead = tuple(pt.shape[1:-2])
    B = int(np.prod(lead)) if lead else 1
    out = torch.empty(lead + (H, W), dtype=dtype, device="cuda")
    _lib.call("jubik_slice_patches_adjoint", dev.code(dtype), dev.ptr(pt), B, H, W, n, m,
              dev.ptr(out), dev.stream())
    return dev.like_input(out, patches)


def convolve(x, y, domain, axes):
    """FFT convolution, jubik/convolve.py:148-176.

    ``x`` is the kernel, ``y`` the signal; returns dvol * Re(ifftn(fftn(x)
    fftn(y))) over ``axes``, the circular convolution.  One or two axes of
    any lengths up to 2048; arrays of equal rank may convolve over any two
    of their axes (they are moved to the end on the device and back).  When
    y has one more axis than x, x is broadcast over it (the last two axes
    only) and the reference's notice is printed.
    """
    nd_x, nd_y = len(x.shape), len(y.shape)
    ax = tuple(a % nd_y for a in axes)
    if len(set(ax)) != len(ax) or not 1 <= len(ax) <= 2:
        raise NotImplementedError("convolve runs over one or two distinct axes")
    dvol = float(math.prod(domain.distances[a] for a in axes))
    broadcast = nd_y > nd_x
    if broadcast:
        print("kernel_shape:", tuple(x.shape))
        print("signal_shape:", tuple(y.shape))
        print("Dimension Inconsistency. Broadcasting PSFs")
        if ax != (nd_y - 2, nd_y - 1):
            raise NotImplementedError("the broadcast branch convolves over the last two axes")
    dev.require_device()
    dtype = dev.float_dtype(x, y)
    xt, yt = dev.to_device(x, dtype), dev.to_device(y, dtype)
    if broadcast:
        H, W = int(yt.shape[-2]), int(yt.shape[-1])
        if tuple(xt.shape[-2:]) != (H, W):
            raise ValueError(f"kernel {tuple(xt.shape)} and signal {tuple(yt.shape)} differ on "
                             "the convolution axes")
        if tuple(yt.shape[:-3]) != tuple(xt.shape[:-2]):
            xt = xt.unsqueeze(-3).expand(yt.shape[:-3] + (1, H, W)).contiguous()
            nbx = int(np.prod(yt.shape[:-3])) if yt.dim() > 3 else 1
        else:
            nbx = int(np.prod(xt.shape[:-2])) if xt.dim() > 2 else 1
        back = None
    else:
        shape = torch.broadcast_shapes(tuple(xt.shape), tuple(yt.shape))
        xt, yt = xt.expand(shape), yt.expand(shape)
        if len(ax) == 1:  # a unit axis in front of the convolution axis
            xt, yt = xt.unsqueeze(-1), yt.unsqueeze(-1)
            ax = (len(shape), ax[0])
        back = ax
        xt = torch.movedim(xt, ax, (-2, -1)).contiguous()
        yt = torch.movedim(yt, ax, (-2, -1)).contiguous()
        H, W = int(yt.shape[-2]), int(yt.shape[-1])
        nbx = int(np.prod(xt.shape[:-2])) if xt.dim() > 2 else 1
    nby = int(np.prod(yt.shape[:-2])) if yt.dim() > 2 else 1
    out = torch.empty(tuple(yt.shape), dtype=dtype, device="cuda")
    if out.numel():
        nbytes = int(_lib.load().jubik_convolve_work_bytes(dev.code(dtype), nby, H, W))
        if nbytes < 0:
            raise NotImplementedError(f"convolve axes of lengths {H} x {W}: at most 2048")
        work = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        _lib.call("jubik_convolve", dev.code(dtype), dev.ptr(xt), dev.ptr(yt), nbx, nby, H, W,
                  dvol, dev.ptr(work), dev.ptr(out), dev.stream())
    if back is not None:
        out = torch.movedim(out, (-2, -1), back)
        if len(axes) == 1:
            out = out.squeeze(-1)
        out = out.contiguous()
    return dev.like_input(out, y)


def integrate(x, domain, axes):
    """dvol * sum(x) over ``axes`` (jubik/convolve.py:179-184).

    The reference evaluates this with numpy on the host at set-up time
    (PSF normalisation, chandra_response.py:123); it is not on the
    per-evaluation path and stays a host reduction here.
    """
    total = np.sum(np.asarray(x), axis=tuple(axes))
    dvol = float(math.prod(domain.distances[i] for i in axes))
    return dvol * total
