"""Gradients with respect to the sky's amplitude fields and offsets
(jubik_model_grad_hyper) against the oracle, whose formulas are pinned by
central differences in tests/test_oracle_math.py."""
from dataclasses import replace

import numpy as np
import pytest
import torch

from oracle import model as om
from tests.problems import make_problem, to_product

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


CASES = [
    dict(H=64, n=4, margin=8, E=1, M=1, seed=91, N=128),     # cropped grid, one energy
    dict(H=64, n=4, margin=8, E=3, M=2, seed=92),            # spectral, two modules
    dict(H=256, n=2, margin=64, E=2, M=1, seed=93, N=1024),  # L = 384, GP grid 1024
]


@pytest.mark.parametrize("cfg", CASES)
@pytest.mark.parametrize("dtype,tol,otol", [(torch.float64, 1e-10, 1e-10),
                                            (torch.float32, 1e-5, 1e-4)])
def test_hyper_gradients_vs_oracle(ja, cfg, dtype, tol, otol):
    prob, xi = make_problem(**cfg)
    want = om.nll_and_grad_hyper(prob, *xi)
    model = to_product(ja, prob, dtype=dtype)
    dev = lambda a: None if a is None else torch.as_tensor(a, device="cuda")
    got = model.nll_and_grad_hyper(dev(xi[0]), dev(xi[1]))
    assert abs(got[0] - want[0]) <= tol * abs(want[0])
    for g, w in zip(got[1:5], want[1:5]):
        assert (g is None) == (w is None)
        if w is not None:
            assert _rel(g.double().cpu().numpy(), w) <= tol
    for g, w in zip(got[5:], want[5:]):
        assert (g is None) == (w is None)
        if w is not None:
            assert abs(g - w) <= otol * abs(w)
    # the position gradients are those of the plain gradient call, bit for bit
    _, g_r, g_a = model.nll_and_grad(dev(xi[0]), dev(xi[1]))
    assert torch.equal(g_r, got[1])
    if g_a is not None:
        assert torch.equal(g_a, got[2])


def test_set_parameters_moves_the_model(ja):
    prob, xi = make_problem(H=64, n=4, margin=8, E=2, M=1, seed=94)
    model = to_product(ja, prob, dtype=torch.float64)
    rng = np.random.default_rng(3)
    amp_ref = prob.amp_ref * (1.0 + 0.1 * rng.random(prob.amp_ref.shape))
    moved = replace(prob, amp_ref=amp_ref, zm=prob.zm + 0.2, mu_alpha=prob.mu_alpha - 0.3)
    model.sky.set_parameters(amp_ref=amp_ref, zm=moved.zm, mu=moved.mu_alpha)
    want = om.nll_and_grad_hyper(moved, *xi)
    dev = lambda a: torch.as_tensor(a, device="cuda")
    got = model.nll_and_grad_hyper(dev(xi[0]), dev(xi[1]))
    assert abs(got[0] - want[0]) <= 1e-10 * abs(want[0])
    assert _rel(got[3].cpu().numpy(), want[3]) <= 1e-10
    assert abs(got[5] - want[5]) <= 1e-10 * abs(want[5])
    with pytest.raises(ValueError):
        model.sky.set_parameters(amp_ref=np.ones((3, 3)))


def test_powerlaw_hyperparameters_from_the_gpu_gradient(ja):
    """The power-law slope and fluctuation gradients chained from the GPU's
    amplitude-field gradient equal those chained from the oracle's."""
    prob, xi = make_problem(H=64, n=4, margin=8, E=1, M=1, seed=95)
    amp = ja.powerlaw_amplitude(64, slope=-3.5, fluctuations=0.8)
    prob = replace(prob, amp_ref=amp)
    want = om.nll_and_grad_hyper(prob, *xi)
    model = to_product(ja, prob, dtype=torch.float64)
    got = model.nll_and_grad_hyper(torch.as_tensor(xi[0], device="cuda"), None)
    w = ja.powerlaw_amplitude_grad(64, want[3], slope=-3.5, fluctuations=0.8)
    g = ja.powerlaw_amplitude_grad(64, got[3], slope=-3.5, fluctuations=0.8)
    assert g[0] == pytest.approx(w[0], rel=1e-10)
    assert g[1] == pytest.approx(w[1], rel=1e-10)
