"""torch.autograd bindings of the HIP path: the backward of the PSF
operator is its adjoint and the backward of the fused NLL is the fused
gradient, checked by torch's numerical Jacobians (gradcheck, FP64) and
against direct library calls."""
import numpy as np
import pytest
import torch

from tests.problems import make_problem, psf_windows, to_product

pytestmark = pytest.mark.gpu


def _op(ja, H, n, m, E, M, dtype, seed):
    rng = np.random.default_rng(seed)
    op = ja.LinpatchOperator((E, H, H), n, m, n_modules=M, dtype=dtype)
    win = np.stack([psf_windows(rng, n, E, 2 * (H // n)) for _ in range(M)])
    op.set_kernel(win if M > 1 else win[0], layout="window")
    return op, rng


@pytest.mark.parametrize("E,M", [(1, 1), (2, 2)])
def test_linpatch_gradcheck(ja, E, M):
    op, rng = _op(ja, 16, 2, 8, E, M, torch.float64, seed=51)
    x = torch.tensor(rng.normal(size=(E, 16, 16)), dtype=torch.float64, device="cuda",
                     requires_grad=True)
    assert torch.autograd.gradcheck(lambda t: ja.linpatch_apply(op, t, 0.7), (x,),
                                    eps=1e-6, atol=1e-9, rtol=1e-6)


def test_linpatch_backward_is_the_adjoint(ja):
    op, rng = _op(ja, 64, 4, 8, 2, 1, torch.float32, seed=52)
    x = torch.tensor(rng.normal(size=(2, 64, 64)), dtype=torch.float32, device="cuda",
                     requires_grad=True)
    y = ja.linpatch_apply(op, x, 1.5)
    (y * y).sum().backward()
    want = op.adjoint(2 * op.forward(x.detach(), 1.5), 1.5)
    assert torch.equal(x.grad, want)


def test_xray_nll_gradcheck(ja):
    prob, xi = make_problem(H=16, n=2, margin=8, E=2, M=1, seed=53)
    model = to_product(ja, prob, dtype=torch.float64)
    xr = torch.tensor(xi[0], dtype=torch.float64, device="cuda", requires_grad=True)
    xa = torch.tensor(xi[1], dtype=torch.float64, device="cuda", requires_grad=True)
    assert torch.autograd.gradcheck(lambda a, b: ja.xray_nll(model, a, b), (xr, xa),
                                    eps=1e-6, atol=1e-3, rtol=1e-5)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_xray_nll_backward_is_the_fused_gradient(ja, dtype):
    prob, xi = make_problem(H=64, n=4, margin=8, E=2, M=2, seed=54)
    model = to_product(ja, prob, dtype=dtype)
    xr = torch.tensor(xi[0], dtype=dtype, device="cuda", requires_grad=True)
    xa = torch.tensor(xi[1], dtype=dtype, device="cuda", requires_grad=True)
    loss = 3.0 * ja.xray_nll(model, xr, xa)
    loss.backward()
    nll, g_r, g_a = model.nll_and_grad(xr.detach(), xa.detach())
    assert float(loss.item()) == 3.0 * nll
    assert torch.equal(xr.grad, (torch.tensor(3.0, dtype=torch.float64, device="cuda")
                                 .to(dtype) * g_r))
    assert torch.equal(xa.grad, (torch.tensor(3.0, dtype=torch.float64, device="cuda")
                                 .to(dtype) * g_a))
