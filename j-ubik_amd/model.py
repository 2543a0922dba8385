"""Fused likelihood of the X-ray forward model: one gradient evaluation per call.

lam_M = Exposure_M(linpatch_convolve(a * s(xi), K_M)) read out where the
counts are valid, NLL = sum(lam - d log lam) (FP64): the likelihood built by
jubik/instruments/erosita/erosita_likelihood.py:51-89 and
jubik/instruments/chandra/chandra_likelihood.py:15-37 composed with the GP
sky.  ``nll_and_grad`` runs forward, Poisson residual and the hand-written
adjoint chain in libjubik_hip.so; ``metric`` gives the MGVI metric product
J^T diag(1/lam) J v at the last gradient position.  ``nll_and_grad_batch``
and ``metric_batch`` evaluate S positions / tangents in one pass (the VI
samples of nifty.re's optimize_kl, demos/x-ray-imaging.py:404-430).
"""
import ctypes

import numpy as np
import torch

from . import _device as dev
from . import _lib
from .convolve import LinpatchOperator
from .sky import CorrelatedFieldSky


def dense_counts(data, keep):
    """Counts as an int32 (M, E, H, W) array with -1 where not read out.

    ``data`` holds one 1-D C-order vector per module (the read-out wire
    format of jubik/response.py:110); ``keep`` is the (M, E, H, W) bool mask.
    """
    keep = np.asarray(keep, dtype=bool)
    out = np.full(keep.shape, -1, dtype=np.int32)
    for i, v in enumerate(data):
        v = np.asarray(v)
        if v.size != int(keep[i].sum()):
            raise ValueError(f"module {i}: {v.size} counts for {int(keep[i].sum())} pixels")
        if v.size and (np.any(v < 0) or np.any(v != np.round(v))):
            raise ValueError("counts must be non-negative integers")
        out[i][keep[i]] = v.astype(np.int32)
    return out


class XrayModel:
    """GP sky + per-module linear-patching PSF + exposure + Poisson data."""

    def __init__(self, sky: CorrelatedFieldSky, psf: LinpatchOperator, pixel_area, exposure,
                 counts, max_samples=1):
        if psf.dtype != sky.dtype:
            raise ValueError("sky and PSF operators must share a dtype")
        if not psf.x_shared:
            raise ValueError("the PSF operator must read one sky for all modules")
        self.sky, self.psf = sky, psf
        self.dtype = sky.dtype
        self.pixel_area = float(pixel_area)
        J = psf.M * psf.E
        shape = (psf.M, psf.E, psf.H, psf.W)
        self.exposure = dev.to_device(exposure, self.dtype).reshape(shape).contiguous()
        self.counts = torch.as_tensor(np.asarray(counts, dtype=np.int32) if not isinstance(
            counts, torch.Tensor) else counts.to(torch.int32), device="cuda").reshape(
                shape).contiguous()
        self._model = ctypes.c_void_p()
        self.max_samples = int(max_samples)
        _lib.call("jubik_model_create_batched", ctypes.byref(self._model), sky._plan, psf._plan,
                  self.pixel_area, self.max_samples)
        _lib.call("jubik_model_set_data", self._model, dev.ptr(self.exposure),
                  dev.ptr(self.counts))
        self._nll = torch.empty(1, dtype=torch.float64, device="cuda")
        self.J = J
        self.col_div, self.row_div = 0, 0
        self._point_version = None  # sky version of the metric base point
        self._batch_S = 0           # samples of the last batched gradient
        self._batch_version = None
        self._solver = None

    def __del__(self):
        m = getattr(self, "_model", None)
        if m is not None and m.value:
            try:
                _lib.load().jubik_model_destroy(m)
            except Exception:
                pass
            self._model = None

    def autotune(self, samples=1, reps=10, cache=None):
        """Pick the fastest column- and row-pass tiles for passes of
        ``samples`` positions (1: single evaluations) by timing one PSF
        forward + adjoint per candidate; the model then uses them for all its
        passes.  ``cache``: a JSON file of earlier choices, read first and
        updated."""
        from .convolve import tile_cache_get, tile_cache_put, tune_result
        key = "model/" + self.psf.tune_key(samples)
        hit = tile_cache_get(cache, key)
        if hit is not None:
            self.set_tile(*hit)
            return {"col_div": hit[0], "row_div": hit[1], "ms": None, "cached": True}
        col, row = ctypes.c_int(), ctypes.c_int()
        ms = (ctypes.c_double * 6)()
        _lib.call("jubik_model_autotune", self._model, int(samples), int(reps),
                  ctypes.byref(col), ctypes.byref(row), ms, dev.stream())
        self.col_div, self.row_div = int(col.value), int(row.value)
        tile_cache_put(cache, key, (self.col_div, self.row_div))
        return tune_result(self.col_div, self.row_div, ms)

    def set_tile(self, col_div, row_div=0):
        """Column- and row-pass tile divisors of the model's passes (0: the
        PSF operator's); results agree to rounding."""
        _lib.call("jubik_model_set_tile", self._model, int(col_div), int(row_div))
        self.col_div, self.row_div = int(col_div), int(row_div)
        return self

    def _grads(self):
        N = self.sky.N
        g_ref = torch.empty((N, N), dtype=self.dtype, device="cuda")
        g_alpha = torch.empty_like(g_ref) if self.sky.has_alpha else None
        return g_ref, g_alpha

    def nll_and_grad_device(self, xi_ref, xi_alpha=None, out=None):
        """Enqueue one gradient evaluation; returns device tensors
        (nll float64 [1], g_ref, g_alpha) without synchronising."""
        s = self.sky
        xr = dev.to_device(xi_ref, self.dtype)
        xa = dev.to_device(xi_alpha, self.dtype) if s.has_alpha else None
        g_ref, g_alpha = out if out is not None else self._grads()
        _lib.call("jubik_model_grad", self._model, dev.ptr(s.amp_ref), dev.ptr(s.amp_alpha),
                  dev.ptr(xr), dev.ptr(xa), s.zm, s.mu, dev.ptr(s.dnu_dev), dev.ptr(self._nll),
                  dev.ptr(g_ref), dev.ptr(g_alpha), dev.stream())
        self._point_version = s.version
        return self._nll, g_ref, g_alpha

    # ------------------------------------------------------------ point sources
    def set_point_sources(self, a, q, alpha_mean=0.0, alpha_std=1.0):
        """Add point sources to the diffuse sky (jubik/sky_models.py:376-441,
        spectral_product_mf_point_sources.py:99-287): per pixel p_e =
        I(xi_ps) exp(alpha dnu_e) with I = invgamma_prior(a, q)(xi_ps) and
        alpha = alpha_mean + alpha_std xi_ps_alpha.  ``a=None`` removes them."""
        if a is None:
            _lib.call("jubik_model_set_point_sources", self._model, 0, 0.0, 0.0, 0.0, 1.0)
            self.point_sources = None
            return self
        _lib.call("jubik_model_set_point_sources", self._model, 1, float(a), float(q),
                  float(alpha_mean), float(alpha_std))
        self.point_sources = dict(a=float(a), q=float(q), alpha_mean=float(alpha_mean),
                                  alpha_std=float(alpha_std))
        self._point_version = None
        return self

    def nll_and_grad_points(self, xi_ref, xi_alpha, xi_ps, xi_ps_alpha=None):
        """NLL and gradients with point sources: (nll, g_ref, g_alpha, g_ps,
        g_ps_alpha).  xi_ps (H, W) evaluates one position and sets the metric
        base point; a stack (S, H, W) with (S, N, N) latents evaluates S
        positions in one pass (nll then an array of S)."""
        if getattr(self, "point_sources", None) is None:
            raise RuntimeError("no point-source prior set (call set_point_sources)")
        s = self.sky
        H, W = self.psf.H, self.psf.W
        xp = dev.to_device(xi_ps, self.dtype)
        batch = xp.dim() == 3
        S = int(xp.shape[0]) if batch else 1
        if batch:
            S = self._batch_size(xp)
            xr = self._stack(xi_ref, S, "xi_ref")
            xa = self._stack(xi_alpha, S, "xi_alpha") if s.has_alpha else None
        else:
            xr = dev.to_device(xi_ref, self.dtype)
            xa = dev.to_device(xi_alpha, self.dtype) if s.has_alpha else None
        want = (S, H, W) if batch else (H, W)
        if tuple(xp.shape) != want:
            raise ValueError(f"xi_ps must have shape {want}, got {tuple(xp.shape)}")
        xpa = None
        if s.has_alpha:
            if xi_ps_alpha is None:
                raise ValueError("a spectral sky needs xi_ps_alpha")
            xpa = dev.to_device(xi_ps_alpha, self.dtype)
            if tuple(xpa.shape) != want:
                raise ValueError(f"xi_ps_alpha must have shape {want}")
            xpa = xpa.contiguous()
        xp = xp.contiguous()
        gshape = (S, s.N, s.N) if batch else (s.N, s.N)
        g_ref = torch.empty(gshape, dtype=self.dtype, device="cuda")
        g_alpha = torch.empty_like(g_ref) if s.has_alpha else None
        g_ps = torch.empty(want, dtype=self.dtype, device="cuda")
        g_psa = torch.empty_like(g_ps) if s.has_alpha else None
        nll = torch.empty(S, dtype=torch.float64, device="cuda")
        _lib.call("jubik_model_grad_points", self._model, S, dev.ptr(s.amp_ref),
                  dev.ptr(s.amp_alpha), dev.ptr(xr), dev.ptr(xa), s.zm, s.mu,
                  dev.ptr(s.dnu_dev), dev.ptr(xp), dev.ptr(xpa), dev.ptr(nll), dev.ptr(g_ref),
                  dev.ptr(g_alpha), dev.ptr(g_ps), dev.ptr(g_psa), dev.stream())
        self.check_domain()
        if batch:
            self._batch_S = S
            self._batch_version = s.version
            out_nll = nll.cpu().numpy()
        else:
            self._point_version = s.version
            out_nll = float(nll.item())
        like = lambda t: None if t is None else dev.like_input(t, xi_ps)
        return out_nll, like(g_ref), like(g_alpha), like(g_ps), like(g_psa)

    def check_domain(self, raise_error=True):
        """Counts of read-out pixels with lam <= 0 per sample of the last
        gradient evaluation (synchronises).  Raises FloatingPointError when
        any is positive (the NLL is then inf or NaN, as in the reference:
        zero-exposure pixels must be masked) unless ``raise_error`` is
        False."""
        n = max(self.max_samples, 1)
        counts = (ctypes.c_int64 * n)()
        st = _lib.load().jubik_model_check_domain(self._model, counts, dev.stream())
        if st != _lib.EDOMAIN or raise_error:
            _lib.check(st)
        return counts

    def nll_and_grad(self, xi_ref, xi_alpha=None):
        """(NLL as a Python float, grad_xi_ref, grad_xi_alpha).  Raises
        FloatingPointError when lam <= 0 at a read-out pixel."""
        nll, g_ref, g_alpha = self.nll_and_grad_device(xi_ref, xi_alpha)
        self.check_domain()
        val = float(nll.item())
        return val, dev.like_input(g_ref, xi_ref), (
            dev.like_input(g_alpha, xi_ref) if g_alpha is not None else None)

    def _check_point(self):
        if self._point_version is None:
            raise RuntimeError("no base point (evaluate the gradient first)")
        if self._point_version != self.sky.version:
            raise RuntimeError("the sky parameters changed since the last gradient evaluation: "
                               "evaluate the gradient at the new parameters first")

    def nll_and_grad_hyper(self, xi_ref, xi_alpha=None):
        """One gradient evaluation that also returns the gradients with
        respect to the sky's amplitude fields and offsets:
        (NLL, grad_xi_ref, grad_xi_alpha, grad_amp_ref, grad_amp_alpha,
        grad_zm, grad_mu); the alpha entries are None for a single-energy
        sky.  These are what the reference's amplitude models and zero-mode
        priors chain their hyperparameter gradients through."""
        s = self.sky
        xr = dev.to_device(xi_ref, self.dtype)
        xa = dev.to_device(xi_alpha, self.dtype) if s.has_alpha else None
        g_ref, g_alpha = self._grads()
        ga_ref, ga_alpha = self._grads()
        off = torch.empty(2, dtype=self.dtype, device="cuda")
        _lib.call("jubik_model_grad_hyper", self._model, dev.ptr(s.amp_ref),
                  dev.ptr(s.amp_alpha), dev.ptr(xr), dev.ptr(xa), s.zm, s.mu,
                  dev.ptr(s.dnu_dev), dev.ptr(self._nll), dev.ptr(g_ref), dev.ptr(g_alpha),
                  dev.ptr(ga_ref), dev.ptr(ga_alpha), dev.ptr(off), dev.stream())
        self._point_version = s.version
        self.check_domain()
        val = float(self._nll.item())
        offs = off.double().cpu().numpy()
        like = lambda t: None if t is None else dev.like_input(t, xi_ref)
        return (val, like(g_ref), like(g_alpha), like(ga_ref), like(ga_alpha), float(offs[0]),
                float(offs[1]) if s.has_alpha else None)

    def metric(self, v_ref, v_alpha=None):
        """J^T diag(1/lam) J v at the position of the last gradient call."""
        self._check_point()
        s = self.sky
        vr = dev.to_device(v_ref, self.dtype)
        va = dev.to_device(v_alpha, self.dtype) if s.has_alpha else None
        o_ref, o_alpha = self._grads()
        _lib.call("jubik_model_metric", self._model, dev.ptr(s.amp_ref), dev.ptr(s.amp_alpha),
                  dev.ptr(s.dnu_dev), dev.ptr(vr), dev.ptr(va), dev.ptr(o_ref),
                  dev.ptr(o_alpha), dev.stream())
        return dev.like_input(o_ref, v_ref), (
            dev.like_input(o_alpha, v_ref) if o_alpha is not None else None)

    def _stack(self, a, S, name):
        N = self.sky.N
        t = dev.to_device(a, self.dtype)
        if tuple(t.shape) != (S, N, N):
            raise ValueError(f"{name} must have shape ({S}, {N}, {N}), got {tuple(t.shape)}")
        return t.contiguous()

    def _batch_size(self, a):
        S = int(a.shape[0])
        if not 1 <= S <= self.max_samples:
            raise ValueError(f"batch of {S} samples outside [1, max_samples={self.max_samples}]")
        return S

    def nll_and_grad_batch(self, xi_ref, xi_alpha=None):
        """S gradient evaluations: xi of shape (S, N, N); returns (nll float64
        numpy (S,), g_ref (S, N, N), g_alpha (S, N, N) or None).  The metric
        base point is not changed."""
        s = self.sky
        S = self._batch_size(xi_ref)
        xr = self._stack(xi_ref, S, "xi_ref")
        xa = self._stack(xi_alpha, S, "xi_alpha") if s.has_alpha else None
        g_ref = torch.empty((S, s.N, s.N), dtype=self.dtype, device="cuda")
        g_alpha = torch.empty_like(g_ref) if s.has_alpha else None
        nll = torch.empty(S, dtype=torch.float64, device="cuda")
        _lib.call("jubik_model_grad_batch", self._model, S, dev.ptr(s.amp_ref),
                  dev.ptr(s.amp_alpha), dev.ptr(xr), dev.ptr(xa), s.zm, s.mu, dev.ptr(s.dnu_dev),
                  dev.ptr(nll), dev.ptr(g_ref), dev.ptr(g_alpha), dev.stream())
        self.check_domain()
        self._batch_S = S
        self._batch_version = s.version
        return nll.cpu().numpy(), dev.like_input(g_ref, xi_ref), (
            dev.like_input(g_alpha, xi_ref) if g_alpha is not None else None)

    def metric_batch(self, v_ref, v_alpha=None):
        """S metric products J^T diag(1/lam) J v at the last gradient position;
        v of shape (S, N, N)."""
        self._check_point()
        s = self.sky
        S = self._batch_size(v_ref)
        vr = self._stack(v_ref, S, "v_ref")
        va = self._stack(v_alpha, S, "v_alpha") if s.has_alpha else None
        o_ref = torch.empty((S, s.N, s.N), dtype=self.dtype, device="cuda")
        o_alpha = torch.empty_like(o_ref) if s.has_alpha else None
        _lib.call("jubik_model_metric_batch", self._model, S, dev.ptr(s.amp_ref),
                  dev.ptr(s.amp_alpha), dev.ptr(s.dnu_dev), dev.ptr(vr), dev.ptr(va),
                  dev.ptr(o_ref), dev.ptr(o_alpha), dev.stream())
        return dev.like_input(o_ref, v_ref), (
            dev.like_input(o_alpha, v_ref) if o_alpha is not None else None)

    def sqrt_metric_batch(self, noise):
        """J^T diag(lam^-1/2) noise at the last gradient position for S
        data-space noise vectors of shape (S, M, E, H, W) (entries where the
        counts are masked are ignored); returns (S, N, N) stacks."""
        self._check_point()
        s = self.sky
        S = self._batch_size(noise)
        shape = (S,) + tuple(self.exposure.shape)
        nz = dev.to_device(noise, self.dtype)
        if tuple(nz.shape) != shape:
            raise ValueError(f"noise must have shape {shape}, got {tuple(nz.shape)}")
        o_ref = torch.empty((S, s.N, s.N), dtype=self.dtype, device="cuda")
        o_alpha = torch.empty_like(o_ref) if s.has_alpha else None
        _lib.call("jubik_model_sqrt_metric_batch", self._model, S, dev.ptr(s.amp_ref),
                  dev.ptr(s.amp_alpha), dev.ptr(s.dnu_dev), dev.ptr(nz), dev.ptr(o_ref),
                  dev.ptr(o_alpha), dev.stream())
        return dev.like_input(o_ref, noise), (
            dev.like_input(o_alpha, noise) if o_alpha is not None else None)

    # ------------------------------------------------------------ solvers
    def reserve_solver(self):
        """Allocate the device-resident CG workspace (set-up call)."""
        _lib.call("jubik_model_reserve_solver", self._model)
        return self

    def metric_samples(self, v_ref, v_alpha=None, samples=None):
        """sum_s J_s^T diag(1/lam_s) J_s v for one tangent (N, N) at the S
        sample positions of the last nll_and_grad_batch call."""
        self._check_samples()
        s = self.sky
        S = int(samples if samples is not None else self._batch_S)
        vr = dev.to_device(v_ref, self.dtype)
        va = dev.to_device(v_alpha, self.dtype) if s.has_alpha else None
        o_ref, o_alpha = self._grads()
        _lib.call("jubik_model_metric_samples", self._model, S, dev.ptr(s.amp_ref),
                  dev.ptr(s.amp_alpha), dev.ptr(s.dnu_dev), dev.ptr(vr), dev.ptr(va),
                  dev.ptr(o_ref), dev.ptr(o_alpha), dev.stream())
        return dev.like_input(o_ref, v_ref), (
            dev.like_input(o_alpha, v_ref) if o_alpha is not None else None)

    def _solve(self, name, S, b_ref, b_alpha, shape, rtol, absdelta, maxiter, check_every):
        s = self.sky
        if not self._solver:
            self.reserve_solver()
            self._solver = True
        br = dev.to_device(b_ref, self.dtype).contiguous()
        ba = dev.to_device(b_alpha, self.dtype).contiguous() if s.has_alpha else None
        for b in (br, ba):
            if b is not None and tuple(b.shape) != shape:
                raise ValueError(f"right-hand sides must have shape {shape}")
        x_ref = torch.empty_like(br)
        x_alpha = torch.empty_like(ba) if ba is not None else None
        nsys = shape[0] if len(shape) == 3 else 1
        info = torch.zeros(2 * nsys, dtype=torch.float64, device="cuda")
        _lib.call(name, self._model, S, dev.ptr(s.amp_ref), dev.ptr(s.amp_alpha),
                  dev.ptr(s.dnu_dev), dev.ptr(br), dev.ptr(ba), dev.ptr(x_ref), dev.ptr(x_alpha),
                  int(maxiter), float(rtol), float(absdelta or 0.0), int(check_every),
                  dev.ptr(info), dev.stream())
        return (x_ref, x_alpha), info

    def metric_solve(self, b_ref, b_alpha=None, rtol=1e-6, absdelta=None, maxiter=100,
                     check_every=10):
        """x_s = (I + J^T diag(1/lam) J)^-1 b_s at the base point for (S, N, N)
        right-hand sides, by the device-resident CG; returns ((x_ref,
        x_alpha), info) with info the per-system iteration counts and
        relative residuals (numpy)."""
        self._check_point()
        N = self.sky.N
        S = int(b_ref.shape[0])
        self._batch_size(b_ref)
        (xr, xa), info = self._solve("jubik_model_metric_solve", S, b_ref, b_alpha, (S, N, N),
                                     rtol, absdelta, maxiter, check_every)
        inf = info.cpu().numpy().reshape(S, 2)
        return (xr, xa), {"iterations": inf[:, 0].astype(np.int64), "rel_residual": inf[:, 1]}

    def _check_samples(self):
        if not self._batch_S:
            raise RuntimeError("no sample positions (evaluate nll_and_grad_batch first)")
        if self._batch_version != self.sky.version:
            raise RuntimeError("the sky parameters changed since the last batched gradient")

    def kl_solve(self, b_ref, b_alpha=None, rtol=1e-6, absdelta=None, maxiter=100,
                 check_every=10):
        """x = (I + (1/S) sum_s J_s^T diag(1/lam_s) J_s)^-1 b at the S sample
        positions of the last nll_and_grad_batch (the KL metric of the
        standardised latents), by the device-resident CG."""
        self._check_samples()
        N = self.sky.N
        S = int(self._batch_S)
        (xr, xa), info = self._solve("jubik_model_kl_solve", S, b_ref, b_alpha, (N, N), rtol,
                                     absdelta, maxiter, check_every)
        inf = info.cpu().numpy()
        return (xr, xa), {"iterations": int(inf[0]), "rel_residual": float(inf[1])}

    # ------------------------------------------------------------ geoVI
    _GEO_MODES = {"K": 0, "KT": 1, "lh": 2, "metric": 3}

    def geovi_point(self, x_ref, x_alpha=None):
        """Evaluate the model at a sample point x for the geoVI operators
        (jubik_model_geovi_point; the metric base point stays where the last
        nll_and_grad put it).  Returns the NLL at x."""
        self._check_point()
        s = self.sky
        xr = dev.to_device(x_ref, self.dtype).contiguous()
        xa = dev.to_device(x_alpha, self.dtype).contiguous() if s.has_alpha else None
        nll = torch.empty(1, dtype=torch.float64, device="cuda")
        _lib.call("jubik_model_geovi_point", self._model, dev.ptr(s.amp_ref),
                  dev.ptr(s.amp_alpha), dev.ptr(xr), dev.ptr(xa), s.zm, s.mu,
                  dev.ptr(s.dnu_dev), dev.ptr(nll), dev.stream())
        self._batch_S = 0  # the batch slot now holds x
        self.check_domain()
        return float(nll.item())

    def geovi_apply(self, mode, v_ref=None, v_alpha=None):
        """The geoVI operators at the base point p and the last geovi_point x:
        "K": J(p)^T D_p D_x J(x) v, "KT": its transpose, "lh": J(p)^T D_p 2
        (sqrt lam(x) - sqrt lam(p)) (no v), "metric": (K + K^T + K^T K) v."""
        s = self.sky
        m = self._GEO_MODES[mode]
        vr = dev.to_device(v_ref, self.dtype).contiguous() if v_ref is not None else None
        va = (dev.to_device(v_alpha, self.dtype).contiguous()
              if (s.has_alpha and v_alpha is not None) else None)
        o_ref, o_alpha = self._grads()
        _lib.call("jubik_model_geovi_apply", self._model, m, dev.ptr(s.amp_ref),
                  dev.ptr(s.amp_alpha), dev.ptr(s.dnu_dev), dev.ptr(vr), dev.ptr(va),
                  dev.ptr(o_ref), dev.ptr(o_alpha), dev.stream())
        return o_ref, o_alpha

    def geovi_solve(self, b_ref, b_alpha=None, rtol=1e-6, absdelta=None, maxiter=100,
                    check_every=10):
        """x = (I + K + K^T + K^T K)^-1 b (the geoVI Gauss-Newton system at the
        last geovi_point) by the device-resident CG; returns ((x_ref,
        x_alpha), info)."""
        s = self.sky
        if not self._solver:
            self.reserve_solver()
            self._solver = True
        br = dev.to_device(b_ref, self.dtype).contiguous()
        ba = dev.to_device(b_alpha, self.dtype).contiguous() if s.has_alpha else None
        x_ref = torch.empty_like(br)
        x_alpha = torch.empty_like(ba) if ba is not None else None
        info = torch.zeros(2, dtype=torch.float64, device="cuda")
        _lib.call("jubik_model_geovi_solve", self._model, dev.ptr(s.amp_ref),
                  dev.ptr(s.amp_alpha), dev.ptr(s.dnu_dev), dev.ptr(br), dev.ptr(ba),
                  dev.ptr(x_ref), dev.ptr(x_alpha), int(maxiter), float(rtol),
                  float(absdelta or 0.0), int(check_every), dev.ptr(info), dev.stream())
        inf = info.cpu().numpy()
        return (x_ref, x_alpha), {"iterations": int(inf[0]), "rel_residual": float(inf[1])}

    def capture_grad(self, samples=0):
        """Capture one gradient evaluation into an executable HIP graph.

        samples = 0 captures the single evaluation (its launches set the
        metric base point), samples = S >= 1 the batched one.  Write new
        positions into ``graph.xi_ref`` / ``graph.xi_alpha`` and call
        ``graph.launch()``; results appear in ``graph.nll``, ``graph.g_ref``
        and ``graph.g_alpha``.
        """
        return GradGraph(self, int(samples))


class _GradSlot:
    """Fixed device buffers of one captured evaluation: positions in,
    NLL and gradients out."""

    def __init__(self, model, samples):
        s = model.sky
        if not 0 <= samples <= model.max_samples:
            raise ValueError(f"samples must be in [0, max_samples={model.max_samples}]")
        shape = ((samples,) if samples else ()) + (s.N, s.N)
        self.model, self.samples = model, samples
        self.xi_ref = torch.zeros(shape, dtype=model.dtype, device="cuda")
        self.xi_alpha = torch.zeros(shape, dtype=model.dtype, device="cuda") if s.has_alpha else None
        self.g_ref = torch.empty(shape, dtype=model.dtype, device="cuda")
        self.g_alpha = torch.empty(shape, dtype=model.dtype, device="cuda") if s.has_alpha else None
        self.nll = torch.empty(max(samples, 1), dtype=torch.float64, device="cuda")
        self.offsets_version = s.offsets_version

    def check_current(self):
        """The captured offsets (zm, mu) are launch arguments fixed at
        capture: refuse to launch after they changed."""
        if self.model.sky.offsets_version != self.offsets_version:
            raise RuntimeError("the sky offsets (zm, mu) changed since this graph was captured: "
                               "capture it again")

    def desc(self):
        s = self.model.sky
        return _lib.GradDesc(self.model._model.value, self.samples, dev.ptr(s.amp_ref),
                             dev.ptr(s.amp_alpha), dev.ptr(self.xi_ref), dev.ptr(self.xi_alpha),
                             s.zm, s.mu, dev.ptr(s.dnu_dev), dev.ptr(self.nll),
                             dev.ptr(self.g_ref), dev.ptr(self.g_alpha))


def capture_grad_batch(entries):
    """Capture the gradient evaluations of several models (observations on
    uneven grids, one model each) into one HIP graph whose evaluations run
    as parallel branches.  ``entries``: (model, samples) pairs, samples = 0
    for a single evaluation (its launches set that model's metric base
    point) or S >= 1 for a batch.  Returns a ``BatchGradGraph``: write
    positions into ``graph.slots[i].xi_ref`` / ``xi_alpha``, call
    ``graph.launch()``, read ``slots[i].nll`` / ``g_ref`` / ``g_alpha``."""
    return BatchGradGraph([_GradSlot(m, int(S)) for m, S in entries])


class _MetricSlot:
    """Fixed device buffers of one captured metric batch: S tangents in,
    S products out."""

    def __init__(self, model, samples):
        s = model.sky
        model._check_point()
        if not 1 <= samples <= model.max_samples:
            raise ValueError(f"samples must be in [1, max_samples={model.max_samples}]")
        shape = (samples, s.N, s.N)
        self.model, self.samples = model, samples
        self.v_ref = torch.zeros(shape, dtype=model.dtype, device="cuda")
        self.v_alpha = torch.zeros(shape, dtype=model.dtype, device="cuda") if s.has_alpha else None
        self.out_ref = torch.empty(shape, dtype=model.dtype, device="cuda")
        self.out_alpha = torch.empty(shape, dtype=model.dtype, device="cuda") if s.has_alpha else None
        self.version = s.version

    def desc(self):
        s = self.model.sky
        return _lib.MetricDesc(self.model._model.value, self.samples, dev.ptr(s.amp_ref),
                               dev.ptr(s.amp_alpha), dev.ptr(s.dnu_dev), dev.ptr(self.v_ref),
                               dev.ptr(self.v_alpha), dev.ptr(self.out_ref),
                               dev.ptr(self.out_alpha))


def capture_metric_batch(entries):
    """Capture the metric products of several models (one per observation,
    uneven grids allowed) into one HIP graph whose batches run as parallel
    branches: ``entries`` are (model, S) pairs, each model at its current
    base point.  Write tangents into ``graph.slots[i].v_ref`` / ``v_alpha``,
    call ``graph.launch()``, read ``out_ref`` / ``out_alpha``."""
    return MetricBatchGraph([_MetricSlot(m, int(S)) for m, S in entries])


class MetricBatchGraph:
    """Several captured metric batches in one executable graph."""

    def __init__(self, slots):
        self.slots = slots
        descs = (_lib.MetricDesc * len(slots))(*[sl.desc() for sl in slots])
        self._graph = ctypes.c_void_p()
        _lib.call("jubik_graph_capture_metric_batch", ctypes.byref(self._graph),
                  ctypes.cast(descs, ctypes.c_void_p), len(slots))

    def launch(self):
        """Enqueue one replay on the current stream (no synchronisation)."""
        for sl in self.slots:
            if sl.model.sky.version != sl.version:
                raise RuntimeError("the sky parameters changed since this graph was captured: "
                                   "capture it again")
        _lib.call("jubik_graph_launch", self._graph, dev.stream())

    def __del__(self):
        g = getattr(self, "_graph", None)
        if g is not None and g.value:
            try:
                _lib.load().jubik_graph_destroy(g)
            except Exception:
                pass
            self._graph = None


class BatchGradGraph:
    """Several captured gradient evaluations in one executable graph."""

    def __init__(self, slots):
        self.slots = slots
        descs = (_lib.GradDesc * len(slots))(*[sl.desc() for sl in slots])
        self._graph = ctypes.c_void_p()
        _lib.call("jubik_graph_capture_batch", ctypes.byref(self._graph),
                  ctypes.cast(descs, ctypes.c_void_p), len(slots))

    def launch(self):
        """Enqueue one replay on the current stream (no synchronisation)."""
        for sl in self.slots:
            sl.check_current()
        _lib.call("jubik_graph_launch", self._graph, dev.stream())
        for sl in self.slots:
            if sl.samples == 0:
                sl.model._point_version = sl.model.sky.version

    def __del__(self):
        g = getattr(self, "_graph", None)
        if g is not None and g.value:
            try:
                _lib.load().jubik_graph_destroy(g)
            except Exception:
                pass
            self._graph = None


class GradGraph:
    """A captured gradient evaluation with its fixed device buffers."""

    def __init__(self, model: XrayModel, samples):
        s = model.sky
        slot = _GradSlot(model, samples)
        self._slot = slot
        self.model = model
        self.xi_ref, self.xi_alpha = slot.xi_ref, slot.xi_alpha
        self.g_ref, self.g_alpha, self.nll = slot.g_ref, slot.g_alpha, slot.nll
        self._graph = ctypes.c_void_p()
        _lib.call("jubik_model_grad_graph", ctypes.byref(self._graph), model._model, samples,
                  dev.ptr(s.amp_ref), dev.ptr(s.amp_alpha), dev.ptr(self.xi_ref),
                  dev.ptr(self.xi_alpha), s.zm, s.mu, dev.ptr(s.dnu_dev), dev.ptr(self.nll),
                  dev.ptr(self.g_ref), dev.ptr(self.g_alpha))

    def launch(self):
        """Enqueue one replay on the current stream (no synchronisation)."""
        self._slot.check_current()
        _lib.call("jubik_graph_launch", self._graph, dev.stream())
        if self._slot.samples == 0:
            self.model._point_version = self.model.sky.version

    def __del__(self):
        g = getattr(self, "_graph", None)
        if g is not None and g.value:
            try:
                _lib.load().jubik_graph_destroy(g)
            except Exception:
                pass
            self._graph = None
