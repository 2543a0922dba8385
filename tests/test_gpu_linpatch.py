"""linpatch_convolve on the GPU against the float64 oracle.

FP64 kernels must agree with the oracle to round-off (rel 1e-10 of the
output's max); FP32 kernels to rel 2e-6 (norm-wise).  Geometries cover the
FFT lengths in use, clamped windows (odd patch stride), n not dividing H,
the energy broadcast of one kernel per patch, and several modules.
"""
import numpy as np
import pytest
import torch

from oracle import convolve as cv
from oracle.data import Domain as ODomain
from tests.problems import psf_windows

pytestmark = pytest.mark.gpu

# (H, n, margin) -> L = 2 * (H // n) + 2 * margin
GEOMS = [
    (64, 4, 8),     # L = 48
    (128, 4, 16),   # L = 96 (single-energy CPU config)
    (100, 4, 7),    # odd stride 25, clamped last window, L = 64
    (102, 4, 7),    # n does not divide H, L = 64
    (96, 3, 16),    # L = 96, three patches per axis
    (256, 8, 32),   # L = 128
    (192, 3, 32),   # L = 192
    (128, 4, 48),   # L = 160? -> not compiled: skipped below
    (512, 8, 64),   # L = 256
    (256, 2, 64),   # L = 384
    (1280, 10, 64), # L = 384, W > 1024: 8 output columns per thread
    (8, 2, 4),      # L = 16, the shortest compiled length
    (16, 2, 8),     # L = 32
    (128, 2, 48),   # L = 224 = 2^5 7
    (162, 2, 81),   # L = 324 = 2^2 3^4, odd stride 81
    (384, 2, 0),    # placeholder, margin 0 handled separately
]


def _cases():
    out = []
    for H, n, m in GEOMS:
        L = 2 * (H // n) + 2 * m
        if m == 0 or L not in (16, 24, 32, 48, 64, 96, 128, 192, 224, 256, 324, 384):
            continue
        out.append((H, n, m))
    return out


def _inputs(H, n, E, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(E, H, H)) + 2.0
    win = psf_windows(rng, n, E, 2 * (H // n))
    return x, win


@pytest.mark.parametrize("H,n,m", _cases())
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_forward_matches_oracle(ja, H, n, m, dtype):
    E = 2
    x, win = _inputs(H, n, E, seed=H + n + m)
    want = cv.linpatch_convolve_window(x, ODomain((E, H, H), (1.0, 1.0, 1.0)), win, n, m)
    op = ja.LinpatchOperator((E, H, H), n, m, dtype=dtype)
    op.set_kernel(win, layout="window")
    got = op.forward(torch.as_tensor(x, device="cuda")).double().cpu().numpy()
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    tol = 1e-12 if dtype == torch.float64 else 2e-6
    print(f"H={H} n={n} m={m} L={op.L} {dtype}: rel err {err:.2e}")
    assert err <= tol


@pytest.mark.parametrize("H,n,m", _cases()[:5] + [(512, 8, 64), (256, 2, 64), (1280, 10, 64),
                                                  (8, 2, 4), (16, 2, 8), (128, 2, 48),
                                                  (162, 2, 81)])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_adjoint_matches_oracle(ja, H, n, m, dtype):
    E = 2
    x, win = _inputs(H, n, E, seed=7 * H + m)
    g = np.random.default_rng(3).normal(size=(E, H, H))
    dom = ODomain((E, H, H), (1.0, 1.0, 1.0))
    want = cv.linpatch_adjoint_window(g, dom, win, n, m)
    op = ja.LinpatchOperator((E, H, H), n, m, dtype=dtype)
    op.set_kernel(win, layout="window")
    got = op.adjoint(torch.as_tensor(g, device="cuda")).double().cpu().numpy()
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    tol = 1e-12 if dtype == torch.float64 else 2e-6
    print(f"adjoint H={H} n={n} m={m}: rel err {err:.2e}")
    assert err <= tol


def test_full_frame_kernel_drop_in(ja):
    """The reference signature with full-frame kernels (float64)."""
    H, n, m, E = 64, 4, 8, 3
    x, win = _inputs(H, n, E, seed=5)
    full = cv.window_to_full_frame(win, (H, H))
    want = cv.linpatch_convolve(x, ODomain((E, H, H), (0.5, 0.5, 0.5)), full, n, m)
    got = ja.linpatch_convolve(x, ja.Domain((E, H, H), (0.5, 0.5, 0.5)), full, n, m)
    assert isinstance(got, np.ndarray) and got.shape == (E, H, H)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())


def test_kernel_broadcast_over_energies(ja, capsys):
    H, n, m, E = 64, 4, 8, 2
    x, win = _inputs(H, n, 1, seed=9)
    x = np.concatenate([x, 2 * x])
    full = cv.window_to_full_frame(win[:, 0], (H, H))  # (P, H, W)
    want = cv.linpatch_convolve(x, ODomain((E, H, H), (1, 1, 1)), full, n, m)
    got = ja.linpatch_convolve(x, ja.Domain((E, H, H), (1, 1, 1)), full, n, m)
    out = capsys.readouterr().out
    assert "Dimension Inconsistency. Broadcasting PSFs" in out
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())


def test_modules_share_input(ja):
    H, n, m, E, M = 64, 4, 8, 2, 3
    rng = np.random.default_rng(21)
    x = rng.normal(size=(E, H, H))
    wins = np.stack([psf_windows(rng, n, E, 2 * (H // n)) for _ in range(M)])
    dom = ODomain((E, H, H), (1, 1, 1))
    want = np.stack([cv.linpatch_convolve_window(x, dom, wins[i], n, m) for i in range(M)])
    full = cv.window_to_full_frame(wins, (H, H))
    got = ja.linpatch_convolve_modules(x, ja.Domain((E, H, H), (1, 1, 1)), full, n, m)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())
    # adjoint sums over modules
    op = ja.LinpatchOperator((E, H, H), n, m, n_modules=M, dtype=torch.float64)
    op.set_kernel(wins, layout="window")
    g = rng.normal(size=(M, E, H, H))
    want_adj = sum(cv.linpatch_adjoint_window(g[i], dom, wins[i], n, m) for i in range(M))
    got_adj = op.adjoint(torch.as_tensor(g, device="cuda")).cpu().numpy()
    np.testing.assert_allclose(got_adj, want_adj, rtol=0, atol=1e-12 * np.abs(want_adj).max())


def test_margin_zero_returns_empty_like_reference(ja):
    H, n = 32, 4
    x, win = _inputs(H, n, 1, seed=1)
    full = cv.window_to_full_frame(win, (H, H))
    got = ja.linpatch_convolve(x, ja.Domain((1, H, H), (1, 1, 1)), full, n, 0)
    assert got.shape == (1, 0, 0)


def test_run_twice_bitwise_equal(ja):
    H, n, m, E = 128, 4, 16, 2
    x, win = _inputs(H, n, E, seed=4)
    op = ja.LinpatchOperator((E, H, H), n, m, dtype=torch.float32)
    op.set_kernel(win, layout="window")
    xt = torch.as_tensor(x, dtype=torch.float32, device="cuda")
    a = op.forward(xt).cpu().numpy()
    b = op.forward(xt).cpu().numpy()
    assert np.array_equal(a, b)
    ga = op.adjoint(torch.as_tensor(a, device="cuda")).cpu().numpy()
    gb = op.adjoint(torch.as_tensor(a, device="cuda")).cpu().numpy()
    assert np.array_equal(ga, gb)


def test_full_size_adjointness_fp32(ja):
    """1024^2 x 4 energies, n = 8, margin 64 (L = 384): <y, A x> = <A^T y, x>."""
    H, n, m, E = 1024, 8, 64, 4
    rng = np.random.default_rng(8)
    win = psf_windows(rng, n, E, 2 * (H // n), moffat=2.5)
    op = ja.LinpatchOperator((E, H, H), n, m, dtype=torch.float32)
    op.set_kernel(win, layout="window")
    x = torch.as_tensor(rng.random((E, H, H)), dtype=torch.float32, device="cuda")
    y = torch.as_tensor(rng.random((E, H, H)), dtype=torch.float32, device="cuda")
    Ax = op.forward(x).double()
    Aty = op.adjoint(y).double()
    lhs = float((y.double() * Ax).sum())
    rhs = float((Aty * x.double()).sum())
    print(f"adjointness at L=384: {lhs:.10e} {rhs:.10e}")
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)
    # linearity
    Ax2 = op.forward(2.0 * x).double()
    assert float((Ax2 - 2 * Ax).abs().max()) <= 1e-5 * float(Ax.abs().max())
