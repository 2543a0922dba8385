"""Batched VI samples: S gradient evaluations / metric products per pass.

A batch must reproduce S single calls bit for bit (the kernels run the same
arithmetic per job), must leave the metric base point alone, and is checked
against the float64 oracle at one sample position.
"""
import numpy as np
import pytest
import torch

from oracle import model as om
from tests.problems import make_problem, to_product

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def _positions(xi, S, seed):
    rng = np.random.default_rng(seed)
    xr = np.stack([xi[0] + 0.05 * rng.normal(size=xi[0].shape) for _ in range(S)])
    xa = None if xi[1] is None else np.stack(
        [xi[1] + 0.05 * rng.normal(size=xi[1].shape) for _ in range(S)])
    return xr, xa


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("E,M", [(1, 1), (2, 2)])
def test_grad_batch_equals_single_calls(ja, dtype, E, M):
    prob, xi = make_problem(H=64, n=4, margin=8, E=E, M=M, seed=11)
    S = 3
    model = to_product(ja, prob, dtype, max_samples=S)
    xr, xa = _positions(xi, S, 1)
    nll_b, gr_b, ga_b = model.nll_and_grad_batch(xr, xa)
    for s in range(S):
        nll, gr, ga = model.nll_and_grad(xr[s], None if xa is None else xa[s])
        assert nll_b[s] == nll
        assert np.array_equal(gr_b[s], gr)
        if xa is not None:
            assert np.array_equal(ga_b[s], ga)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_metric_batch_equals_single_calls(ja, dtype):
    prob, xi = make_problem(H=64, n=4, margin=8, E=2, M=2, seed=12)
    S = 4
    model = to_product(ja, prob, dtype, max_samples=S)
    model.nll_and_grad(xi[0], xi[1])
    rng = np.random.default_rng(2)
    vr = rng.normal(size=(S,) + xi[0].shape)
    va = rng.normal(size=(S,) + xi[1].shape)
    single = [model.metric(vr[s], va[s]) for s in range(S)]
    # a batched gradient in between must not move the base point
    model.nll_and_grad_batch(*_positions(xi, 2, 3))
    br, ba = model.metric_batch(vr, va)
    for s in range(S):
        assert np.array_equal(br[s], single[s][0])
        assert np.array_equal(ba[s], single[s][1])


def test_grad_batch_vs_oracle_fp64(ja):
    prob, xi = make_problem(H=64, n=4, margin=8, E=2, M=1, seed=13)
    S = 2
    model = to_product(ja, prob, torch.float64, max_samples=S)
    xr, xa = _positions(xi, S, 4)
    nll_b, gr_b, ga_b = model.nll_and_grad_batch(xr, xa)
    for s in range(S):
        nll, (gr, ga) = om.nll_and_grad(prob, xr[s], xa[s])
        assert abs(nll_b[s] - nll) <= 1e-12 * abs(nll)
        assert _rel(gr_b[s], gr) <= 1e-11
        assert _rel(ga_b[s], ga) <= 1e-11


def test_antithetic_metric_batch(ja):
    """32 metric products (16 antithetic pairs +-v) on the 256^2 observation
    (n = 6, margin 22, L = 128) in one pass: the pairs cancel exactly and a
    pair checks against the oracle."""
    prob, xi = make_problem(H=256, n=6, margin=22, E=1, M=1, seed=4)
    S = 32
    model = to_product(ja, prob, torch.float32, max_samples=S)
    model.nll_and_grad(xi[0])
    rng = np.random.default_rng(44)
    v = rng.normal(size=(S // 2,) + xi[0].shape)
    vs = np.concatenate([v, -v])
    out, _ = model.metric_batch(torch.as_tensor(vs, device="cuda"))
    out = out.double().cpu().numpy()
    assert np.array_equal(out[: S // 2], -out[S // 2:])
    for s in (0, 7):
        want, _ = om.metric_product(prob, xi[0], None, v[s])
        assert _rel(out[s], want) <= 1e-5


def test_batch_size_is_checked(ja):
    prob, xi = make_problem(H=32, n=2, margin=8, E=1, M=1, seed=14)
    model = to_product(ja, prob, torch.float32, max_samples=2)
    with pytest.raises(ValueError, match="max_samples"):
        model.nll_and_grad_batch(np.stack([xi[0]] * 3))
    with pytest.raises(RuntimeError, match="base point"):
        model.metric_batch(np.stack([xi[0]] * 2))
