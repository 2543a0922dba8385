"""The split-BF16 matrix-core column transform (set_transform("mfma_bf16x3"))
of FP32 plans at FFT length 384: each 384-point column DFT runs as three
128-point DFTs on v_mfma_f32_32x32x16_bf16 (hi*hi + hi*lo + lo*hi, FP32
accumulation) and an FP32 radix-3 step.

Held to the float64 oracle: the split products keep about 16 mantissa bits
of each factor, so a transform carries relative errors near 4e-6 (a CPU
emulation of the kernel's arithmetic gives 3.7e-6 per 384-point DFT)
against 1e-7 for the FP32 FFT; the tolerances below are written for that."""
import numpy as np
import pytest
import torch

from oracle import convolve as ocv
from oracle import model as om
from oracle.data import Domain
from tests.problems import make_problem, psf_windows, to_product

pytestmark = pytest.mark.gpu

TOL_MFMA = 2e-5      # forward / adjoint, relative 2-norm vs float64
TOL_GRAD = 1e-4      # gradient, relative 2-norm vs float64
TOL_NLL = 1e-6       # NLL, relative


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


# (H, n, margin, E, M): every case has window + 2 margin = 384
GEOMS = [
    (512, 4, 64, 2, 1),   # window 256
    (384, 4, 96, 1, 1),   # window 192, wide margin
    (640, 4, 32, 1, 1),   # window 320, narrow margin
    (520, 4, 62, 1, 1),   # n does not divide H: clamped windows
    (512, 4, 64, 2, 2),   # two modules, one input
]


@pytest.mark.parametrize("H,n,m,E,M", GEOMS)
def test_mfma_transform_vs_oracle(ja, H, n, m, E, M):
    rng = np.random.default_rng(H + n + m + M)
    w = 2 * (H // n)
    assert w + 2 * m == 384
    win = np.stack([psf_windows(rng, n, E, w, sigma=(1.0, 4.0), moffat=2.5) for _ in range(M)])
    op = ja.LinpatchOperator((E, H, H), n, m, n_modules=M, dtype=torch.float32)
    assert op.L == 384
    op.set_kernel(win if M > 1 else win[0], layout="window")
    ref = ja.LinpatchOperator((E, H, H), n, m, n_modules=M, dtype=torch.float32)
    ref.set_kernel(win if M > 1 else win[0], layout="window")
    op.set_transform("mfma_bf16x3")
    dom = Domain((E, H, H), (1.0, 1.0, 1.0))
    x = rng.random((E, H, H))
    g = rng.normal(size=(M, E, H, H) if M > 1 else (E, H, H))
    xt = torch.as_tensor(x, dtype=torch.float32, device="cuda")
    gt = torch.as_tensor(g, dtype=torch.float32, device="cuda")
    got_y = op.forward(xt).double().cpu().numpy()
    got_g = op.adjoint(gt).double().cpu().numpy()
    fft_y = ref.forward(xt).double().cpu().numpy()
    fft_g = ref.adjoint(gt).double().cpu().numpy()
    for mod in range(M):
        want_y = ocv.linpatch_convolve_window(x, dom, win[mod], n, m)
        gy = g[mod] if M > 1 else g
        gy_got = got_y[mod] if M > 1 else got_y
        assert _rel(gy_got, want_y) <= TOL_MFMA, ("forward", mod)
        assert _rel(fft_y[mod] if M > 1 else fft_y, want_y) <= 2e-6
    want_g = sum(ocv.linpatch_adjoint_window(g[mod] if M > 1 else g, dom, win[mod], n, m)
                 for mod in range(M))
    assert _rel(got_g, want_g) <= TOL_MFMA, "adjoint"
    assert _rel(fft_g, want_g) <= 2e-6
    # the matrix-core path is a different rounding, not the same numbers
    assert not np.array_equal(got_y, fft_y)
    # adjointness of the matrix-core pair itself
    lhs = float(np.vdot(got_y.ravel(), g.ravel()))
    rhs = float(np.vdot(x.ravel(), got_g.ravel()))
    assert abs(lhs - rhs) <= 1e-4 * abs(lhs)


def test_switching_back_restores_the_fft_results(ja):
    rng = np.random.default_rng(5)
    H, n, m, E = 512, 4, 64, 1
    win = psf_windows(rng, n, E, 256)
    op = ja.LinpatchOperator((E, H, H), n, m, dtype=torch.float32)
    op.set_kernel(win, layout="window")
    x = torch.as_tensor(rng.random((E, H, H)), dtype=torch.float32, device="cuda")
    a = op.forward(x).clone()
    op.set_transform("mfma_bf16x3")
    b = op.forward(x).clone()
    op.set_transform("fft")
    c = op.forward(x).clone()
    assert torch.equal(a, c)
    assert not torch.equal(a, b)
    # the matrix-core results are deterministic
    op.set_transform("mfma_bf16x3")
    assert torch.equal(op.forward(x), b)


def test_mfma_transform_rejects_other_plans(ja):
    op64 = ja.LinpatchOperator((1, 512, 512), 4, 64, dtype=torch.float64)
    with pytest.raises(NotImplementedError, match="float32"):
        op64.set_transform("mfma_bf16x3")
    op_other = ja.LinpatchOperator((1, 256, 256), 4, 32, dtype=torch.float32)
    with pytest.raises(NotImplementedError, match="length 192"):
        op_other.set_transform("mfma_bf16x3")
    op_fold = ja.LinpatchOperator((1, 306, 306), 3, 64, dtype=torch.float32)
    with pytest.raises(NotImplementedError):
        op_fold.set_transform("mfma_bf16x3")
    with pytest.raises(ValueError):
        op_other.set_transform("bf16")
    # a rejected request leaves the plan on its FFT
    assert op_other.transform == "fft"


@pytest.mark.parametrize("cfg", [dict(H=512, n=4, margin=64, E=2, M=1, seed=91, N=512),
                                 dict(H=512, n=4, margin=64, E=1, M=2, seed=92, N=512,
                                      moffat=2.5)])
def test_mfma_gradient_vs_oracle(ja, cfg):
    prob, xi = make_problem(**cfg)
    want_nll, (wr, wa) = om.nll_and_grad(prob, *xi)
    model = to_product(ja, prob, torch.float32)
    nll_f, gr_f, ga_f = model.nll_and_grad(*xi)
    model.psf.set_transform("mfma_bf16x3")
    nll, gr, ga = model.nll_and_grad(*xi)
    e_mfma = _rel(gr, wr)
    e_fft = _rel(gr_f, wr)
    print(f"gradient rel. error vs float64: fft {e_fft:.3e}, mfma_bf16x3 {e_mfma:.3e}")
    assert abs(nll - want_nll) <= TOL_NLL * abs(want_nll)
    assert e_mfma <= TOL_GRAD
    if wa is not None:
        assert _rel(ga, wa) <= TOL_GRAD


def test_mfma_batched_gradients_and_metric(ja):
    """S samples per pass (the sample loop of the persistent kernel and the
    per-sample NLL finalisation) against S single evaluations."""
    prob, xi = make_problem(H=512, n=4, margin=64, E=2, M=1, seed=93, N=512)
    model = to_product(ja, prob, torch.float32, max_samples=3)
    model.psf.set_transform("mfma_bf16x3")
    rng = np.random.default_rng(3)
    xr = np.stack([xi[0] + 0.05 * rng.normal(size=xi[0].shape) for _ in range(3)])
    xa = np.stack([xi[1] + 0.05 * rng.normal(size=xi[1].shape) for _ in range(3)])
    nlls, gr, ga = model.nll_and_grad_batch(xr, xa)
    for s in range(3):
        n1, g1, a1 = model.nll_and_grad(xr[s], xa[s])
        assert nlls[s] == pytest.approx(n1, rel=1e-12)
        assert _rel(gr[s], g1) <= 1e-6
        assert _rel(ga[s], a1) <= 1e-6
    # metric products at the last single point against the FFT path's
    v = rng.normal(size=(3,) + xi[0].shape)
    va = rng.normal(size=(3,) + xi[0].shape)
    m_r, m_a = model.metric_batch(v, va)
    model.psf.set_transform("fft")
    model.nll_and_grad(xr[2], xa[2])
    f_r, f_a = model.metric_batch(v, va)
    assert _rel(m_r, f_r) <= TOL_GRAD
    assert _rel(m_a, f_a) <= TOL_GRAD
