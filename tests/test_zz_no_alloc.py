"""jubik_poisson_nll_grad_noalloc through the C ABI: the same NLL and
gradient, bit for bit, as jubik_poisson_nll_grad with a caller-provided
workspace, for two calls in a row on one GPU (the second with a smaller
input, so stale partials of the first call in the shared scratch would
show), in float32 and float64."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


@pytest.mark.parametrize("dtype,code", [(torch.float32, 0), (torch.float64, 1)])
def test_noalloc_poisson_matches_workspace_entry_point(ja, dtype, code):
    from jubik_amd import _lib
    lib = _lib.load()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rng = np.random.default_rng(8)
    work = torch.empty(_lib.POISSON_WORK_BYTES // 8, dtype=torch.float64, device="cuda")
    for n in (1_000_003, 40_001):  # many reduction blocks, then fewer
        lam = torch.as_tensor(rng.uniform(0.2, 30.0, size=n), dtype=dtype, device="cuda")
        d = torch.as_tensor(rng.poisson(lam.double().cpu().numpy()), dtype=dtype, device="cuda")
        nll_a = torch.empty(1, dtype=torch.float64, device="cuda")
        nll_b = torch.empty(1, dtype=torch.float64, device="cuda")
        g_a, g_b = torch.empty_like(lam), torch.empty_like(lam)
        _lib.check(lib.jubik_poisson_nll_grad(code, _ptr(lam), _ptr(d), n, _ptr(nll_a),
                                              _ptr(g_a), _ptr(work), stream))
        _lib.check(lib.jubik_poisson_nll_grad_noalloc(code, _ptr(lam), _ptr(d), n, _ptr(nll_b),
                                                      _ptr(g_b), stream))
        torch.cuda.synchronize()
        assert float(nll_a.item()) == float(nll_b.item())
        assert torch.equal(g_a, g_b)
        want = float(np.sum(lam.double().cpu().numpy() -
                            d.double().cpu().numpy() * np.log(lam.double().cpu().numpy())))
        assert abs(float(nll_b.item()) - want) <= 1e-9 * abs(want)


def test_noalloc_poisson_validates_arguments(ja):
    from jubik_amd import _lib
    lib = _lib.load()
    assert lib.jubik_poisson_nll_grad_noalloc(2, None, None, 0, None, None, None) == _lib.EINVAL
    assert lib.jubik_poisson_nll_grad_noalloc(0, None, None, -1, None, None, None) == _lib.EINVAL
