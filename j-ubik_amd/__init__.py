"""jubik_amd: J-UBIK's X-ray forward model and gradient on AMD MI355X (gfx950).

Drop-in names (jubik/__init__.py:31-59): Domain, linpatch_convolve,
integrate, build_exposure_function, build_readout_function; plus the
Poisson likelihood, the GP sky and the fused gradient / metric model.  All
array work runs in libjubik_hip.so (hand-written HIP kernels behind a C ABI,
include/jubik_hip.h); there is no CPU fallback.
"""
from .data import Domain, Vector
from .convolve import (LinpatchOperator, convolve, integrate, linpatch_convolve,
                       linpatch_convolve_modules, slice_patches, slice_patches_transpose)
from .response import build_exposure_function, build_readout_function
from .likelihood import Poissonian
from .sky import CorrelatedFieldSky, powerlaw_amplitude, powerlaw_amplitude_grad
from .model import XrayModel, capture_grad_batch, capture_metric_batch, dense_counts
from .sampling import conjugate_gradient, draw_linear_samples, metric_solve
from .kl import kl_value_and_grad, newton_cg, optimize_kl
from .geovi import draw_nonlinear_samples, nonlinear_update
from .mock import create_mock_counts, poisson_sample, save_mock_data
from .psf import psf_interpolator
from .dft import DFT128
from . import geovi, priors, synthetic
from .autograd import linpatch_apply, xray_nll

__all__ = [
    "Domain", "Vector", "LinpatchOperator", "convolve", "integrate", "linpatch_convolve",
    "slice_patches", "slice_patches_transpose",
    "linpatch_convolve_modules", "build_exposure_function", "build_readout_function",
    "Poissonian", "CorrelatedFieldSky", "powerlaw_amplitude", "XrayModel", "dense_counts",
    "conjugate_gradient", "draw_linear_samples", "metric_solve", "kl_value_and_grad",
    "newton_cg", "optimize_kl", "draw_nonlinear_samples", "nonlinear_update", "create_mock_counts", "save_mock_data",
    "poisson_sample", "psf_interpolator", "DFT128",
]
