"""Regression vectors of the float64 oracle (tests/golden/, self-generated
by tests/golden/make_golden.py): the oracle must keep reproducing them, and
the GPU kernels must match them."""
import ast
from pathlib import Path

import numpy as np
import pytest

from oracle import model as om
from tests.problems import make_problem

GOLDEN = Path(__file__).resolve().parent / "golden"
CASES = sorted(GOLDEN.glob("*.npz"))


def _load(path):
    z = np.load(path)  # allow_pickle stays False
    cfg = ast.literal_eval(str(z["cfg"]))
    return z, cfg


@pytest.mark.parametrize("path", CASES, ids=[p.stem for p in CASES])
def test_oracle_reproduces_golden(path):
    z, cfg = _load(path)
    prob, xi = make_problem(**cfg)
    np.testing.assert_array_equal(xi[0], z["xi_ref"])
    nll, (gr, ga) = om.nll_and_grad(prob, *xi)
    assert nll == pytest.approx(float(z["nll"]), rel=1e-13)
    np.testing.assert_allclose(gr, z["g_ref"], rtol=1e-11, atol=1e-11 * np.abs(z["g_ref"]).max())
    if "g_alpha" in z:
        np.testing.assert_allclose(ga, z["g_alpha"], rtol=1e-11,
                                   atol=1e-11 * np.abs(z["g_alpha"]).max())


@pytest.mark.gpu
@pytest.mark.parametrize("path", CASES, ids=[p.stem for p in CASES])
def test_kernels_match_golden(ja, path):
    import torch
    from tests.problems import to_product
    z, cfg = _load(path)
    prob, xi = make_problem(**cfg)
    for dtype, tol in ((torch.float64, 1e-10), (torch.float32, 1e-5)):
        model = to_product(ja, prob, dtype)
        nll, gr, ga = model.nll_and_grad(*[None if v is None else torch.as_tensor(v, device="cuda")
                                           for v in xi])
        got = np.concatenate([gr.double().cpu().numpy().ravel()] +
                             ([ga.double().cpu().numpy().ravel()] if ga is not None else []))
        want = np.concatenate([z["g_ref"].ravel()] +
                              ([z["g_alpha"].ravel()] if "g_alpha" in z else []))
        assert np.linalg.norm(got - want) <= tol * np.linalg.norm(want)
        assert abs(nll - float(z["nll"])) <= tol * abs(float(z["nll"]))
        v = [torch.as_tensor(z["v_ref"], device="cuda")]
        if "v_alpha" in z:
            v.append(torch.as_tensor(z["v_alpha"], device="cuda"))
        mr, ma = model.metric(*v)
        got = np.concatenate([mr.double().cpu().numpy().ravel()] +
                             ([ma.double().cpu().numpy().ravel()] if ma is not None else []))
        want = np.concatenate([z["m_ref"].ravel()] +
                              ([z["m_alpha"].ravel()] if "m_alpha" in z else []))
        assert np.linalg.norm(got - want) <= tol * np.linalg.norm(want)
