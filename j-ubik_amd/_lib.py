"""ctypes binding of libjubik_hip.so (the C ABI declared in include/jubik_hip.h).

The library is the only compute path: if it is missing or no HIP device is
visible, every compute call raises.  Status codes map to the exceptions the
reference raises (ValueError for invalid arguments).
"""
import ctypes as C
import os
import threading
from pathlib import Path

LIB_PATH = Path(os.environ.get("JUBIK_HIP_LIB") or
                Path(__file__).resolve().parent / "libjubik_hip.so")

OK, EINVAL, EUNSUPPORTED, EHIP, ENOMEM, ENOTREADY, EDOMAIN = 0, -1, -2, -3, -4, -5, -6
FLOAT32, FLOAT64 = 0, 1
KERNEL_FULL, KERNEL_WINDOW = 0, 1
POISSON_WORK_BYTES = 8192

_vp, _i64, _int, _dbl = C.c_void_p, C.c_int64, C.c_int, C.c_double
_dblp = C.POINTER(C.c_double)

# name -> (argtypes); every function returns int except jubik_last_error.
SIGNATURES = {
    "jubik_abi_version": [],
    "jubik_timing_enable": [_int],
    "jubik_timing_report": [C.c_char_p, _i64],
    "jubik_timing_select": [C.c_char_p],
    "jubik_debug_fail_alloc": [_i64],
    "jubik_linpatch_create": [C.POINTER(_vp), _int, _i64, _i64, _i64, _i64, _i64, _i64, _int,
                              _int],
    "jubik_linpatch_destroy": [_vp],
    "jubik_linpatch_info": [_vp, C.POINTER(_i64)],
    "jubik_linpatch_set_kernel": [_vp, _vp, _int, _vp],
    "jubik_linpatch_forward": [_vp, _vp, _vp, _dbl, _vp],
    "jubik_linpatch_adjoint": [_vp, _vp, _vp, _dbl, _vp],
    "jubik_slice_patches": [_int, _vp, _i64, _i64, _i64, _i64, _i64, _vp, _vp],
    "jubik_slice_patches_adjoint": [_int, _vp, _i64, _i64, _i64, _i64, _i64, _vp, _vp],
    "jubik_convolve": [_int, _vp, _vp, _i64, _i64, _i64, _i64, _dbl, _vp, _vp, _vp],
    "jubik_exposure_apply": [_int, _vp, _i64, _vp, _i64, _vp, _i64, _vp],
    "jubik_readout_gather": [_int, _vp, _vp, _i64, _vp, _vp],
    "jubik_readout_scatter": [_int, _vp, _vp, _i64, _i64, _vp, _vp],
    "jubik_poisson_nll_grad": [_int, _vp, _vp, _i64, _vp, _vp, _vp, _vp],
    "jubik_poisson_metric": [_int, _vp, _vp, _i64, _int, _vp, _vp],
    "jubik_poisson_nll_grad_noalloc": [_int, _vp, _vp, _i64, _vp, _vp, _vp],
    "jubik_gp_create": [C.POINTER(_vp), _int, _i64, _i64, _i64, _i64, _int],
    "jubik_gp_destroy": [_vp],
    "jubik_gp_forward": [_vp, _vp, _vp, _vp, _vp, _dbl, _dbl, _vp, _vp, _vp],
    "jubik_gp_vjp": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "jubik_model_create": [C.POINTER(_vp), _vp, _vp, _dbl],
    "jubik_model_destroy": [_vp],
    "jubik_model_set_data": [_vp, _vp, _vp],
    "jubik_model_grad": [_vp, _vp, _vp, _vp, _vp, _dbl, _dbl, _vp, _vp, _vp, _vp, _vp],
    "jubik_model_set_point_sources": [_vp, _int, _dbl, _dbl, _dbl, _dbl],
    "jubik_model_grad_points": [_vp, _i64, _vp, _vp, _vp, _vp, _dbl, _dbl, _vp, _vp, _vp, _vp,
                                _vp, _vp, _vp, _vp, _vp],
    "jubik_invgamma_prior": [_int, _vp, _i64, _dbl, _dbl, _vp, _vp, _vp],
    "jubik_model_geovi_point": [_vp, _vp, _vp, _vp, _vp, _dbl, _dbl, _vp, _vp, _vp],
    "jubik_model_geovi_apply": [_vp, _int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "jubik_model_geovi_solve": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _int, _dbl, _dbl, _int,
                                _vp, _vp],
    "jubik_model_grad_hyper": [_vp, _vp, _vp, _vp, _vp, _dbl, _dbl, _vp, _vp, _vp, _vp, _vp, _vp,
                               _vp, _vp],
    "jubik_model_metric": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "jubik_model_check_domain": [_vp, _vp, _vp],
    "jubik_linpatch_set_tile": [_vp, _int, _int],
    "jubik_linpatch_set_transform": [_vp, _int],
    "jubik_linpatch_autotune": [_vp, _int, C.POINTER(_int), C.POINTER(_int), _dblp, _vp],
    "jubik_model_set_tile": [_vp, _int, _int],
    "jubik_model_autotune": [_vp, _i64, _int, C.POINTER(_int), C.POINTER(_int), _dblp, _vp],
    "jubik_model_create_batched": [C.POINTER(_vp), _vp, _vp, _dbl, _i64],
    "jubik_model_grad_batch": [_vp, _i64, _vp, _vp, _vp, _vp, _dbl, _dbl, _vp, _vp, _vp, _vp,
                               _vp],
    "jubik_model_metric_batch": [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "jubik_model_sqrt_metric_batch": [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "jubik_model_metric_samples": [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "jubik_model_reserve_solver": [_vp],
    "jubik_model_metric_solve": [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _int, _dbl, _dbl,
                                 _int, _vp, _vp],
    "jubik_model_kl_solve": [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _int, _dbl, _dbl,
                             _int, _vp, _vp],
    "jubik_poisson_sample": [_int, _vp, _i64, C.c_uint64, _vp, _vp],
    "jubik_model_grad_graph": [C.POINTER(_vp), _vp, _i64, _vp, _vp, _vp, _vp, _dbl, _dbl, _vp,
                               _vp, _vp, _vp],
    "jubik_graph_launch": [_vp, _vp],
    "jubik_graph_capture_batch": [C.POINTER(_vp), _vp, _i64],
    "jubik_graph_capture_metric_batch": [C.POINTER(_vp), _vp, _i64],
    "jubik_graph_destroy": [_vp],
    "jubik_dft128_create": [C.POINTER(_vp)],
    "jubik_dft128_run": [_vp, _int, _vp, _vp, _i64, _vp],
    "jubik_dft128_destroy": [_vp],
    "jubik_psf_interpolate": [_int, _vp, _i64, _i64, _i64, _vp, _vp, _vp, _vp, _i64, _i64, _dbl,
                              _dbl, _i64, _i64, _i64, _int, _vp, _vp],
}



class GradDesc(C.Structure):
    """jubik_grad_desc: one evaluation of a multi-model graph."""
    _fields_ = [("model", _vp), ("samples", _i64), ("amp_ref", _vp), ("amp_alpha", _vp),
                ("xi_ref", _vp), ("xi_alpha", _vp), ("zm", _dbl), ("mu", _dbl), ("dnu", _vp),
                ("nll", _vp), ("g_ref", _vp), ("g_alpha", _vp)]


class MetricDesc(C.Structure):
    """jubik_metric_desc: one metric batch of a multi-model graph."""
    _fields_ = [("model", _vp), ("samples", _i64), ("amp_ref", _vp), ("amp_alpha", _vp),
                ("dnu", _vp), ("v_ref", _vp), ("v_alpha", _vp), ("out_ref", _vp),
                ("out_alpha", _vp)]


_lock = threading.Lock()
_lib = None


def load():
    """Load the shared library once; raise if it has not been built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build the HIP library first "
                "(python -c 'import __graft_entry__ as g; g.build()')")
        lib = C.CDLL(str(LIB_PATH))
        lib.jubik_last_error.restype = C.c_char_p
        lib.jubik_last_error.argtypes = []
        for name, args in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = C.c_int
        lib.jubik_convolve_work_bytes.argtypes = [_int, _i64, _i64, _i64]
        lib.jubik_convolve_work_bytes.restype = _i64
        _lib = lib
        return lib


_EXC = {EINVAL: ValueError, EUNSUPPORTED: NotImplementedError, EHIP: RuntimeError,
        ENOMEM: MemoryError, ENOTREADY: RuntimeError, EDOMAIN: FloatingPointError}


def check(status):
    """Raise the exception that corresponds to a C ABI status code."""
    if status == OK:
        return
    msg = load().jubik_last_error().decode(errors="replace")
    raise _EXC.get(status, RuntimeError)(msg)


def call(name, *args):
    check(getattr(load(), name)(*args))


def timing_enable(on=True):
    call("jubik_timing_enable", int(bool(on)))


def timing_select(name=None):
    """Time only the kernels called ``name`` (None: every kernel)."""
    call("jubik_timing_select", None if name is None else name.encode())


def timing_report():
    """{kernel: (count, total_ms, min_ms, max_ms)} since the last report."""
    buf = C.create_string_buffer(1 << 16)
    call("jubik_timing_report", buf, len(buf))
    out = {}
    for line in buf.value.decode().splitlines():
        name, cnt, tot, lo, hi = line.split()
        out[name] = (int(cnt), float(tot), float(lo), float(hi))
    return out
