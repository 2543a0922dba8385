"""FFT lengths beyond the plain compiled list: the reference demos' geometries.

* L = 320 = 2^6 5 (the Chandra demo: 512^2 sky, 4 x 4 patches, margin
  0.062 * 512 -> 32; demos/configs/chandra_demo.yaml:13,44-45 with
  chandra_response.py:62) runs a radix-5 plan directly.
* Window lengths with a large prime factor, L = 332 = 4 83 (10 x 10 patches
  of stride 102 with the default margin 64) and L = 116 = 4 29 (6 x 6
  patches of stride 42, margin 16), run in fold mode: the transforms use a
  compiled length of at least twice the window, where the size-L circular
  convolution is recovered exactly by folding (LinGeom in common.hpp).
* GP grids 567 = 3^4 7 and 1134 = 2 3^4 7, the padded grids ducc0's
  good_size gives for the demos' 512 and 1024 skies (jubik/sky_models.py:306).

Each case is held to the float64 oracle (scipy FFTs of the exact length)."""
import numpy as np
import pytest
import torch

from oracle import convolve as ocv
from oracle import model as om
from oracle import sky as osky
from oracle.data import Domain
from tests.problems import make_problem, powerlaw_amp, psf_windows, to_product

pytestmark = pytest.mark.gpu

TOL = {torch.float64: 1e-12, torch.float32: 2e-6}


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


# (H, n, margin, E, window length w + 2m, FFT length the plan runs)
GEOMS = [
    (512, 4, 32, 1, 320, 320),    # Chandra demo, radix 5
    (306, 3, 64, 1, 332, 432),    # stride 102, default margin: 332 = 4 * 83, fold
    (256, 6, 16, 1, 116, 192),    # stride 42, margin 16: 116 = 4 * 29, fold
    (64, 4, 1, 2, 34, 64),        # margin 1: heavy aliasing, fold
    (64, 4, 17, 1, 66, 96),       # margin above w / 2 with an uncompiled length
    (70, 4, 5, 2, 44, 96),        # n does not divide H (clamped windows), fold
    (1024, 10, 64, 1, 332, 432),  # the full 10 x 10 geometry, fold
]


@pytest.mark.parametrize("H,n,m,E,Lw,Lfft", GEOMS)
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_linpatch_lengths_vs_oracle(ja, H, n, m, E, Lw, Lfft, dtype):
    rng = np.random.default_rng(H + n + m)
    w = 2 * (H // n)
    assert w + 2 * m == Lw
    win = psf_windows(rng, n, E, w, sigma=(1.0, 4.0), moffat=2.5)
    op = ja.LinpatchOperator((E, H, H), n, m, dtype=dtype)
    assert op.L == Lfft
    op.set_kernel(win, layout="window")
    dom = Domain((E, H, H), (1.0, 1.0, 1.0))
    x = rng.random((E, H, H))
    g = rng.normal(size=(E, H, H))
    want_y = ocv.linpatch_convolve_window(x, dom, win, n, m)
    want_g = ocv.linpatch_adjoint_window(g, dom, win, n, m)
    got_y = op.forward(torch.as_tensor(x, dtype=dtype, device="cuda")).double().cpu().numpy()
    got_g = op.adjoint(torch.as_tensor(g, dtype=dtype, device="cuda")).double().cpu().numpy()
    assert _rel(got_y, want_y) <= TOL[dtype], f"forward, L = {Lw} on {Lfft}"
    assert _rel(got_g, want_g) <= TOL[dtype], f"adjoint, L = {Lw} on {Lfft}"


def test_fold_mode_full_frame_kernel(ja):
    """The reference's full-frame (P, H, W) kernel layout in fold mode."""
    H, n, m = 64, 4, 3
    rng = np.random.default_rng(1)
    w = 2 * (H // n)
    win = psf_windows(rng, n, 1, w)[:, 0]
    full = ocv.window_to_full_frame(win, (H, H))
    op = ja.LinpatchOperator((H, H), n, m, kernel_per_energy=False, dtype=torch.float64)
    op.set_kernel(full, layout="full")
    x = rng.random((H, H))
    want = ocv.linpatch_convolve(x, Domain((H, H), (1.0, 1.0)), full, n, m)
    got = op.forward(torch.as_tensor(x, device="cuda")).cpu().numpy()
    assert _rel(got, want) <= 1e-12


def test_fold_mode_rejects_partial_tiles(ja):
    op = ja.LinpatchOperator((64, 64), 4, 3, kernel_per_energy=False, dtype=torch.float32)
    with pytest.raises(ValueError, match="fold mode"):
        op.set_tile(2, 1)
    op.set_tile(1, 1)


@pytest.mark.parametrize("cfg", [
    dict(H=64, n=4, margin=5, E=2, M=2, seed=41),           # fold mode, two modules
    dict(H=256, n=6, margin=16, E=1, M=1, seed=42),         # L = 116 geometry
    dict(H=512, n=4, margin=32, E=1, M=1, seed=43, N=567),  # the Chandra demo: L = 320, GP 567
])
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-10), (torch.float32, 1e-5)])
def test_model_gradient_at_demo_lengths(ja, cfg, dtype, tol):
    prob, xi = make_problem(**cfg)
    want_nll, (want_r, want_a) = om.nll_and_grad(prob, *xi)
    model = to_product(ja, prob, dtype)
    dev = lambda a: None if a is None else torch.as_tensor(a, dtype=dtype, device="cuda")
    nll, g_r, g_a = model.nll_and_grad(dev(xi[0]), dev(xi[1]))
    assert abs(nll - want_nll) <= tol * abs(want_nll)
    assert _rel(g_r.double().cpu().numpy(), want_r) <= tol
    if want_a is not None:
        assert _rel(g_a.double().cpu().numpy(), want_a) <= tol
    v = np.random.default_rng(7).normal(size=xi[0].shape)
    va = np.random.default_rng(8).normal(size=xi[1].shape) if xi[1] is not None else None
    got_m, _ = model.metric(dev(v), dev(va))
    want_m, _ = om.metric_product(prob, xi[0], xi[1], v, va)
    assert _rel(got_m.double().cpu().numpy(), want_m) <= tol


@pytest.mark.parametrize("N,H,E", [(567, 512, 1), (567, 512, 3), (1134, 1024, 2)])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_gp_demo_grids_vs_oracle(ja, N, H, E, dtype):
    rng = np.random.default_rng(N + E)
    amp = powerlaw_amp(N)
    amp_a = 0.5 * amp if E > 1 else None
    dnu = np.log(np.linspace(1, 3, E)) if E > 1 else None
    xr = rng.normal(size=(N, N))
    xa = rng.normal(size=(N, N)) if E > 1 else None
    want = osky.gp_sky(xr, xa, amp, amp_a, -2.0, -1.0, dnu if E > 1 else [0.0], crop=(H, H))
    sky = ja.CorrelatedFieldSky(N, (H, H), amp, amp_a, zm=-2.0, mu=-1.0, dnu=dnu, dtype=dtype)
    dev = lambda a: None if a is None else torch.as_tensor(a, device="cuda")
    got = sky(dev(xr), dev(xa))
    tol = TOL[dtype] * (1 if dtype == torch.float64 else 1.5)
    assert _rel(got.double().cpu().numpy(), want) <= tol
    cot = rng.normal(size=want.shape)
    g_ref, g_alpha, _, _ = osky.gp_adjoint(want * cot, amp, amp_a, dnu if E > 1 else [0.0],
                                           grid_shape=(N, N))
    gr, ga = sky.vjp(dev(want), dev(cot))
    assert _rel(gr.double().cpu().numpy(), g_ref) <= tol
    if E > 1:
        assert _rel(ga.double().cpu().numpy(), g_alpha) <= tol
