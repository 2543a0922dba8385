"""Error paths under injected allocation failures (jubik_debug_fail_alloc):
every creation call reports JUBIK_ENOMEM with the buffer it was allocating,
leaves no plan behind, frees what it had allocated, and the library works
normally afterwards."""
import ctypes

import numpy as np
import pytest
import torch

from tests.problems import make_problem, to_product

pytestmark = pytest.mark.gpu


def _free():
    """Device memory not held by the library: free plus what torch's caching
    allocator has reserved (torch's own tensors stay out of the count)."""
    torch.cuda.synchronize()
    return torch.cuda.mem_get_info()[0] + torch.cuda.memory_reserved()


def test_injected_allocation_failures_release_everything(ja):
    from jubik_amd import _lib
    lib = _lib.load()
    plan = ctypes.c_void_p()
    before = _free()
    # a linpatch plan makes six allocations (twiddles, tent, spectra, R1, R2, beta)
    for k in range(1, 7):
        _lib.check(lib.jubik_debug_fail_alloc(k))
        st = lib.jubik_linpatch_create(ctypes.byref(plan), 0, 256, 256, 2, 1, 4, 32, 1, 1)
        assert st == _lib.ENOMEM, k
        assert b"injected failure" in lib.jubik_last_error()
        assert not plan.value
    _lib.check(lib.jubik_debug_fail_alloc(0))
    assert _free() == before
    # the same creation succeeds once disarmed
    _lib.check(lib.jubik_linpatch_create(ctypes.byref(plan), 0, 256, 256, 2, 1, 4, 32, 1, 1))
    _lib.check(lib.jubik_linpatch_destroy(plan))
    gp = ctypes.c_void_p()
    for k in range(1, 4):
        _lib.check(lib.jubik_debug_fail_alloc(k))
        assert lib.jubik_gp_create(ctypes.byref(gp), 1, 256, 256, 256, 2, 1) == _lib.ENOMEM
        assert not gp.value
    _lib.check(lib.jubik_debug_fail_alloc(0))
    assert _free() == before


def test_injected_failure_in_model_creation_and_python_surface(ja):
    prob, xi = make_problem(H=64, n=4, margin=8, E=2, M=1, seed=71)
    from jubik_amd import _lib
    lib = _lib.load()
    model = to_product(ja, prob, torch.float32, max_samples=2)
    before = _free()
    for k in (1, 4, 9):
        _lib.check(lib.jubik_debug_fail_alloc(k))
        with pytest.raises(MemoryError, match="injected failure"):
            ja.XrayModel(model.sky, model.psf, prob.pixel_area,
                         np.stack([m.exposure for m in prob.modules]),
                         model.counts.cpu().numpy(), max_samples=3)
    _lib.check(lib.jubik_debug_fail_alloc(0))
    assert _free() == before
    # the existing model is untouched
    nll, g, _ = model.nll_and_grad(xi[0], xi[1])
    assert np.isfinite(nll) and np.all(np.isfinite(g))
