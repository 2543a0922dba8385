"""The fused model (Poisson epilogue, GP-chained adjoint, metric, batches)
on the awkward geometries of the linear patching: odd patch strides with a
clamped last window, n not dividing H with an odd patch count (one patch
column without a partner in the paired transforms), a GP grid cropped to
the sky grid, and one kernel shared by every energy."""
import numpy as np
import pytest
import torch

from oracle import model as om
from tests.problems import make_problem, to_product

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


CASES = [
    dict(H=100, n=4, margin=7, E=2, M=2, seed=61, N=128),   # stride 25 odd: clamped window
    dict(H=98, n=3, margin=16, E=3, M=1, seed=62, N=128),   # n does not divide H, odd n
    dict(H=60, n=5, margin=12, E=1, M=3, seed=63, N=64),    # odd n, three modules, L = 48
    # odd sky width (the GP adjoint reads the folded cotangent element-wise),
    # clamped last window, spectral
    dict(H=99, n=3, margin=15, E=2, M=1, seed=64, N=128),
]


@pytest.mark.parametrize("cfg", CASES)
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-10), (torch.float32, 1e-5)])
def test_grad_and_metric_vs_oracle(ja, cfg, dtype, tol):
    prob, xi = make_problem(**cfg)
    want_nll, (want_r, want_a) = om.nll_and_grad(prob, *xi)
    model = to_product(ja, prob, dtype, max_samples=2)
    nll, g_r, g_a = model.nll_and_grad(*xi)
    got = [g_r] + ([g_a] if g_a is not None else [])
    want = [want_r] + ([want_a] if want_a is not None else [])
    assert abs(nll - want_nll) <= tol * abs(want_nll)
    assert _rel(np.concatenate([np.ravel(g) for g in got]),
                np.concatenate([np.ravel(w) for w in want])) <= tol
    rng = np.random.default_rng(7)
    v = (rng.normal(size=xi[0].shape), None if xi[1] is None else rng.normal(size=xi[1].shape))
    want_m = om.metric_product(prob, xi[0], xi[1], *v)
    got_m = model.metric(*v)
    gm = np.concatenate([np.ravel(g) for g in got_m if g is not None])
    wm = np.concatenate([np.ravel(w) for w in want_m if w is not None])
    assert _rel(gm, wm) <= tol
    # a batch of two positions equals two single evaluations
    xb = np.stack([xi[0], xi[0] + 0.01])
    xab = None if xi[1] is None else np.stack([xi[1], xi[1] - 0.01])
    nb, gb, _ = model.nll_and_grad_batch(xb, xab)
    n1, g1, _ = model.nll_and_grad(xb[1], None if xab is None else xab[1])
    assert nb[1] == n1 and np.array_equal(gb[1], g1)


def test_wide_image_fp32_vs_oracle(ja):
    """W = 1280 (8 output columns per thread in the overlap kernels) on a
    2048 GP grid, FP32: NLL, gradient and a metric product against the
    oracle."""
    prob, xi = make_problem(H=1280, n=10, margin=64, E=1, M=1, seed=66, N=2048)
    want_nll, (want_r, _) = om.nll_and_grad(prob, *xi)
    model = to_product(ja, prob, dtype=torch.float32)
    nll, g_r, _ = model.nll_and_grad(torch.as_tensor(xi[0], device="cuda"))
    assert abs(nll - want_nll) <= 1e-5 * abs(want_nll)
    assert _rel(g_r.double().cpu().numpy(), want_r) <= 1e-5
    v = np.random.default_rng(5).normal(size=xi[0].shape)
    got, _ = model.metric(torch.as_tensor(v, device="cuda"))
    want, _ = om.metric_product(prob, xi[0], None, v, None)
    assert _rel(got.double().cpu().numpy(), want) <= 1e-5


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-10), (torch.float32, 1e-5)])
def test_kernel_shared_by_energies(ja, dtype, tol):
    """kernel_per_energy=False (convolve.py:167-171 broadcast): the oracle
    problem carries the same window for every energy."""
    prob, xi = make_problem(H=64, n=4, margin=8, E=3, M=1, seed=64)
    win0 = prob.modules[0].window[:, :1]
    prob.modules[0].window = np.repeat(win0, 3, axis=1)
    want_nll, (want_r, want_a) = om.nll_and_grad(prob, *xi)
    E, H, W = prob.shape
    sky = ja.CorrelatedFieldSky(64, (H, W), prob.amp_ref, prob.amp_alpha, zm=prob.zm,
                                mu=prob.mu_alpha, dnu=prob.dnu, dtype=dtype)
    psf = ja.LinpatchOperator((E, H, W), prob.n, prob.margin, kernel_per_energy=False, dtype=dtype)
    psf.set_kernel(win0[:, 0], layout="window")
    counts = ja.dense_counts([prob.modules[0].data], prob.modules[0].keep[None])
    model = ja.XrayModel(sky, psf, prob.pixel_area, prob.modules[0].exposure[None], counts)
    nll, g_r, g_a = model.nll_and_grad(*xi)
    assert abs(nll - want_nll) <= tol * abs(want_nll)
    assert _rel(np.concatenate([g_r.ravel(), g_a.ravel()]),
                np.concatenate([want_r.ravel(), want_a.ravel()])) <= tol
