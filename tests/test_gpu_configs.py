"""The geometries of the benchmark configurations on the GPU.

Multi-module: 7 modules x 3 energies, n = 10, margin 60 (L = 324 = 2^2 3^4).
Uneven grids: three observations on 256^2 / 512^2 / 768^2 grids with
    n = 6 / 8 / 12 and margins 22 / 32 / 48 (L = 128 / 192 / 224) and
    metric products.
Wide: 2048^2 x 4 energies, n = 16, margin 64 (L = 384).
Gradient workload: 1024^2 x 4 energies, n = 8, margin 64 (L = 384), spectral
sky, Gaussian+Moffat PSFs, exposure ramp, 5 % mask, against the float64
oracle at full size.
Small instances are checked against the float64 oracle; the other full sizes
through FP32 vs FP64 kernels, adjointness and linearity.
"""
import numpy as np
import pytest
import torch

from oracle import convolve as cv
from oracle import model as om
from oracle.data import Domain as ODomain
from tests.problems import make_problem, psf_windows, to_product

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 2e-6)])
def test_modules_geometry_L324_vs_oracle(ja, dtype, tol):
    """L = 324 on a 303^2 grid (n = 3, stride 101, margin 61): the last
    window start 202 is clamped to H + 2h - w = 201; 2 modules x 3 energies."""
    H, n, m, E, M = 303, 3, 61, 3, 2
    rng = np.random.default_rng(31)
    x = rng.random((E, H, H))
    wins = np.stack([psf_windows(rng, n, E, 2 * (H // n), sigma=(2, 6),
                                 rotation=k * np.pi / 7) for k in range(M)])
    dom = ODomain((E, H, H), (1, 1, 1))
    op = ja.LinpatchOperator((E, H, H), n, m, n_modules=M, dtype=dtype)
    assert op.L == 324
    op.set_kernel(wins, layout="window")
    got = op.forward(torch.as_tensor(x, device="cuda")).double().cpu().numpy()
    want = np.stack([cv.linpatch_convolve_window(x, dom, wins[k], n, m) for k in range(M)])
    assert _rel(got, want) <= tol
    g = rng.normal(size=(M, E, H, H))
    got = op.adjoint(torch.as_tensor(g, device="cuda")).double().cpu().numpy()
    want = sum(cv.linpatch_adjoint_window(g[k], dom, wins[k], n, m) for k in range(M))
    assert _rel(got, want) <= tol


def test_modules_1024x3E_full_size_adjointness(ja):
    """1024^2 x 3 energies, 7 modules, n = 10, margin 60 (FP32)."""
    H, n, m, E, M = 1024, 10, 60, 3, 7
    rng = np.random.default_rng(3)
    wins = np.stack([psf_windows(rng, n, E, 2 * (H // n), sigma=(2, 6),
                                 rotation=k * np.pi / 7) for k in range(M)])
    op = ja.LinpatchOperator((E, H, H), n, m, n_modules=M, dtype=torch.float32)
    op.set_kernel(wins, layout="window")
    x = torch.as_tensor(rng.random((E, H, H)), dtype=torch.float32, device="cuda")
    y = torch.as_tensor(rng.random((M, E, H, H)), dtype=torch.float32, device="cuda")
    lhs = float((y.double() * op.forward(x).double()).sum())
    rhs = float((op.adjoint(y).double() * x.double()).sum())
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


@pytest.mark.parametrize("H,n,m,L", [(256, 6, 22, 128), (512, 8, 32, 192), (768, 12, 48, 224)])
def test_uneven_grids_vs_oracle(ja, H, n, m, L):
    rng = np.random.default_rng(H)
    x = rng.random((1, H, H))
    win = psf_windows(rng, n, 1, 2 * (H // n))
    dom = ODomain((1, H, H), (1, 1, 1))
    want = cv.linpatch_convolve_window(x, dom, win, n, m)
    for dtype, tol in ((torch.float64, 1e-12), (torch.float32, 2e-6)):
        op = ja.LinpatchOperator((1, H, H), n, m, dtype=dtype)
        assert op.L == L
        op.set_kernel(win, layout="window")
        got = op.forward(torch.as_tensor(x, device="cuda")).double().cpu().numpy()
        assert _rel(got, want) <= tol


def test_uneven_grids_metric_products_vs_oracle(ja):
    """Metric products J^T diag(1/lam) J v for several antithetic samples on
    the 256^2 observation (n = 6, margin 22)."""
    prob, xi = make_problem(H=256, n=6, margin=22, E=1, M=1, seed=41, N=256)
    model = to_product(ja, prob, torch.float32)
    model.nll_and_grad(torch.as_tensor(xi[0], device="cuda"))
    rng = np.random.default_rng(4)
    for s in range(2):
        v = rng.normal(size=xi[0].shape)
        for sign in (1.0, -1.0):
            want, _ = om.metric_product(prob, xi[0], None, sign * v)
            got, _ = model.metric(torch.as_tensor(sign * v, device="cuda"))
            assert _rel(got.double().cpu().numpy(), want) <= 1e-5


def test_wide_2048x4E_fp32_vs_fp64(ja):
    """2048^2 x 4 energies, n = 16, margin 64 (1024 patch-energies)."""
    H, n, m, E = 2048, 16, 64, 4
    rng = np.random.default_rng(5)
    win = psf_windows(rng, n, E, 2 * (H // n))
    x = rng.random((E, H, H))
    g = rng.normal(size=(E, H, H))
    out = {}
    for dtype in (torch.float64, torch.float32):
        op = ja.LinpatchOperator((E, H, H), n, m, dtype=dtype)
        op.set_kernel(win, layout="window")
        y = op.forward(torch.as_tensor(x, device="cuda")).double().cpu().numpy()
        gx = op.adjoint(torch.as_tensor(g, device="cuda")).double().cpu().numpy()
        out[dtype] = (y, gx)
        del op
        torch.cuda.empty_cache()
    assert _rel(out[torch.float32][0], out[torch.float64][0]) <= 2e-6
    assert _rel(out[torch.float32][1], out[torch.float64][1]) <= 2e-6


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 2e-6)])
def test_gp_2048_vs_oracle(ja, dtype, tol):
    """The 2048^2 spectral GP and its VJP against the oracle."""
    from oracle import sky as osky
    from tests.problems import powerlaw_amp
    N, E = 2048, 4
    rng = np.random.default_rng(6)
    amp = powerlaw_amp(N)
    dnu = np.log(np.array([1.0, 1.6, 2.5, 4.0]))
    xr, xa = rng.normal(size=(N, N)), rng.normal(size=(N, N))
    want = osky.gp_sky(xr, xa, amp, 0.5 * amp, -2.0, -1.0, dnu)
    sky = ja.CorrelatedFieldSky(N, (N, N), amp, 0.5 * amp, zm=-2.0, mu=-1.0, dnu=dnu,
                                dtype=dtype)
    got = sky(torch.as_tensor(xr, device="cuda"), torch.as_tensor(xa, device="cuda"))
    assert _rel(got.double().cpu().numpy(), want) <= tol
    cot = rng.normal(size=want.shape)
    g_ref, g_alpha, _, _ = osky.gp_adjoint(want * cot, amp, 0.5 * amp, dnu, grid_shape=(N, N))
    gr, ga = sky.vjp(torch.as_tensor(want, device="cuda"), torch.as_tensor(cot, device="cuda"))
    assert _rel(gr.double().cpu().numpy(), g_ref) <= tol
    assert _rel(ga.double().cpu().numpy(), g_alpha) <= tol


def test_wide_2048x4E_gradient_fp32_vs_fp64(ja):
    """The full 2048^2 x 4 energy gradient (GP, PSF with n = 16, margin 64,
    exposure, Poisson) in FP32 against the FP64 build, for one evaluation
    and for a batch of VI samples in one pass."""
    H, n, m, E = 2048, 16, 64, 4
    rng = np.random.default_rng(5)
    from tests.problems import powerlaw_amp
    amp = powerlaw_amp(H)
    dnu = np.log(np.array([1.0, 1.6, 2.5, 4.0]))
    win = psf_windows(rng, n, E, 2 * (H // n), sigma=(1.0, 4.0), moffat=2.5)
    ramp = np.linspace(5e2, 2e3, H)[None, None, :]
    expo = np.broadcast_to(ramp, (E, H, H)).copy()
    expo[rng.random(expo.shape) < 0.05] = 0.0
    keep = expo > 0
    xt, xat = rng.normal(size=(H, H)), rng.normal(size=(H, H))
    S = 3
    xs = [xt[None] + 0.1 * rng.normal(size=(S, H, H)), xat[None] + 0.1 * rng.normal(size=(S, H, H))]

    def build(dtype, counts, S=1):
        sky = ja.CorrelatedFieldSky(H, (H, H), amp, 0.5 * amp, zm=-2.0, mu=-1.0, dnu=dnu,
                                    dtype=dtype)
        psf = ja.LinpatchOperator((E, H, H), n, m, dtype=dtype)
        psf.set_kernel(win, layout="window")
        return sky, psf, ja.XrayModel(sky, psf, 1.0, expo[None], counts[None], max_samples=S)

    sky64, psf64, _ = build(torch.float64, np.zeros(keep.shape, np.int32))
    lam = psf64.forward(sky64(torch.as_tensor(xt, device="cuda"),
                              torch.as_tensor(xat, device="cuda"))) * torch.as_tensor(
        expo, device="cuda")
    counts = np.where(keep, ja.poisson_sample(lam, seed=5).cpu().numpy(), -1).astype(np.int32)
    del sky64, psf64, lam
    res = {}
    for dtype in (torch.float64, torch.float32):
        _, _, model = build(dtype, counts, S)
        nll, gr, ga = model.nll_and_grad_batch(*[torch.as_tensor(v, dtype=dtype, device="cuda")
                                                 for v in xs])
        res[dtype] = (nll, gr.double().cpu().numpy(), ga.double().cpu().numpy())
        del model, gr, ga
        torch.cuda.empty_cache()
    for s in range(S):
        g32 = np.concatenate([res[torch.float32][1][s].ravel(), res[torch.float32][2][s].ravel()])
        g64 = np.concatenate([res[torch.float64][1][s].ravel(), res[torch.float64][2][s].ravel()])
        err = _rel(g32, g64)
        nmax = float(np.max(np.abs(g32 - g64)) / np.max(np.abs(g64)))
        print(f"2048^2 x 4E sample {s}: FP32 vs FP64 grad rel {err:.2e}, normalised max {nmax:.2e}")
        assert err <= 1e-5
        n64 = res[torch.float64][0][s]
        assert abs(res[torch.float32][0][s] - n64) <= 1e-6 * abs(n64)


_WORKLOAD = {}


def _workload_oracle():
    """The 1024^2 x 4E problem and its float64 oracle NLL, gradient and one
    metric product (computed once for both precisions)."""
    if not _WORKLOAD:
        prob, (xr, xa) = make_problem(H=1024, n=8, margin=64, E=4, M=1, seed=75,
                                      expo_range=(5e2, 2e3), mask_frac=0.05, moffat=2.5,
                                      sigma=(1.0, 4.0))
        want_nll, want_g = om.nll_and_grad(prob, xr, xa)
        rng = np.random.default_rng(76)
        v = (rng.normal(size=xr.shape), rng.normal(size=xa.shape))
        _WORKLOAD.update(prob=prob, x=(xr, xa), nll=want_nll, g=want_g, v=v,
                         m=om.metric_product(prob, xr, xa, *v))
    return _WORKLOAD


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-10), (torch.float32, 1e-5)])
def test_gradient_workload_1024x4E_vs_oracle(ja, dtype, tol):
    """The benchmarked evaluation at full size (1024^2 x 4 energies, n = 8,
    margin 64, L = 384, spectral sky, Gaussian+Moffat PSFs of core sigma
    1-4 px, exposure ramp 5e2-2e3 s, 5 % mask): NLL, gradient and one metric
    product against the float64 oracle, the FP32 kernels within the 1e-5
    gradient budget."""
    w = _workload_oracle()
    model = to_product(ja, w["prob"], dtype)
    nll, g_r, g_a = model.nll_and_grad(*w["x"])
    assert abs(nll - w["nll"]) <= tol * abs(w["nll"])
    assert _rel(np.concatenate([np.ravel(g_r), np.ravel(g_a)]),
                np.concatenate([np.ravel(g) for g in w["g"]])) <= tol
    got_m = model.metric(*w["v"])
    assert _rel(np.concatenate([np.ravel(g) for g in got_m]),
                np.concatenate([np.ravel(m) for m in w["m"]])) <= tol
