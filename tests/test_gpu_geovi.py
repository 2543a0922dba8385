"""geoVI nonlinear sample updates on the GPU (jubik_model_geovi_*,
jubik_amd.geovi) against the oracle (oracle/geovi.py, pinned by
tests/test_oracle_geovi.py)."""
import numpy as np
import pytest
import torch

from oracle import geovi as og
from tests.problems import make_problem, to_product

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a = a.double().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def _setup(ja, dtype, E=2, seed=121):
    prob, (pr, pa) = make_problem(H=64, n=4, margin=8, E=E, M=1, seed=seed,
                                  expo_range=(5.0, 20.0))
    rng = np.random.default_rng(seed)
    p = (pr, pa)
    x = (pr + 0.2 * rng.normal(size=pr.shape),
         None if pa is None else pa + 0.2 * rng.normal(size=pa.shape))
    v = (rng.normal(size=pr.shape), None if pa is None else rng.normal(size=pa.shape))
    model = to_product(ja, prob, dtype, max_samples=2)
    return prob, model, p, x, v


@pytest.mark.parametrize("E", [2, 1])
@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-10), (torch.float32, 2e-5)])
def test_geovi_operators_vs_oracle(ja, E, dtype, tol):
    prob, model, p, x, v = _setup(ja, dtype, E=E)
    model.nll_and_grad(*p)
    model.geovi_point(*x)
    got = {"K": model.geovi_apply("K", *v), "KT": model.geovi_apply("KT", *v),
           "lh": model.geovi_apply("lh"), "metric": model.geovi_apply("metric", *v)}
    want = {"K": og.cross(prob, p, x, v, to_p=True), "KT": og.cross(prob, p, x, v, to_p=False),
            "lh": og.lh_part(prob, p, x), "metric": og.gn_metric(prob, p, x, v)}
    for k in got:
        assert _rel(got[k][0], want[k][0]) <= tol, k
        if E > 1:
            assert _rel(got[k][1], want[k][1]) <= tol, k
        else:
            assert want[k][1] is None


def test_geovi_solve_solves_the_gauss_newton_system(ja):
    """Converged on a well-conditioned problem (exposures of 0.05-0.2 s); on
    the problem of the other tests (condition number in the thousands) the
    CG's own residual after a fixed number of iterations equals the oracle's
    |b - A x| / |b| for the returned x."""
    prob, (pr, pa) = make_problem(H=64, n=4, margin=8, E=2, M=1, seed=122,
                                  expo_range=(0.05, 0.2))
    rng = np.random.default_rng(122)
    p = (pr, pa)
    x = (pr + 0.2 * rng.normal(size=pr.shape), pa + 0.2 * rng.normal(size=pa.shape))
    v = (rng.normal(size=pr.shape), rng.normal(size=pa.shape))
    model = to_product(ja, prob, torch.float64, max_samples=2)
    model.nll_and_grad(*p)
    model.geovi_point(*x)
    (sr, sa), info = model.geovi_solve(v[0], v[1], rtol=1e-9, maxiter=500)
    s = (sr.cpu().numpy(), sa.cpu().numpy())
    gm = og.gn_metric(prob, p, x, s)
    assert _rel(s[0] + gm[0], v[0]) <= 1e-7 and _rel(s[1] + gm[1], v[1]) <= 1e-7
    assert info["rel_residual"] <= 1e-9
    prob, model, p, x, v = _setup(ja, torch.float64)
    model.nll_and_grad(*p)
    model.geovi_point(*x)
    (sr, sa), info = model.geovi_solve(v[0], v[1], rtol=1e-12, maxiter=40)
    s = (sr.cpu().numpy(), sa.cpu().numpy())
    gm = og.gn_metric(prob, p, x, s)
    res = np.sqrt(np.sum((v[0] - s[0] - gm[0]) ** 2) + np.sum((v[1] - s[1] - gm[1]) ** 2))
    res /= np.sqrt(np.sum(v[0] ** 2) + np.sum(v[1] ** 2))
    assert info["iterations"] == 40
    assert info["rel_residual"] == pytest.approx(res, rel=1e-5)


def test_nonlinear_update_drives_the_residual_down(ja):
    prob, model, p, x, v = _setup(ja, torch.float64)
    m = (0.5 * v[0], 0.5 * v[1])
    model.nll_and_grad(*p)
    xr, xa, hist = ja.nonlinear_update(model, p[0], p[1], m[0], m[1], x[0], x[1], maxiter=8,
                                       xtol=1e-10, cg_rtol=1e-8, cg_maxiter=100)
    e = hist["energy"]
    assert hist["steps"] >= 2 and all(b < a for a, b in zip(e, e[1:]))
    assert e[-1] <= 1e-6 * e[0]
    # the oracle agrees on the final residual
    r = og.residual(prob, p, m, (xr.cpu().numpy(), xa.cpu().numpy()))
    assert og.half_sq(r) == pytest.approx(e[-1], rel=1e-6, abs=1e-12)
    # the geoVI sample differs from the linear start
    assert float((xr.cpu() - torch.as_tensor(x[0])).abs().max()) > 1e-3


def test_nonlinear_samples_and_optimize_kl(ja):
    prob, model, p, _, _ = _setup(ja, torch.float64)
    sr, sa, info = ja.draw_nonlinear_samples(model, p[0], p[1], n_samples=2, seed=3,
                                             maxiter=4, cg_maxiter=50, draw_maxiter=200)
    assert sr.shape == (2,) + p[0].shape and torch.isfinite(sr).all() and torch.isfinite(sa).all()
    assert len(info["nonlinear"]) == 2
    for h in info["nonlinear"]:
        assert h["energy"][-1] <= h["energy"][0]
    xr, xa, s_ref, s_alpha, hist = ja.optimize_kl(
        model, p[0], p[1], n_samples=2, n_iterations=1, maxiter=2,
        sample_mode="nonlinear_resample", nonlinear_kwargs=dict(maxiter=3, cg_maxiter=40))
    assert torch.isfinite(xr).all() and len(hist) == 1


def test_geovi_requires_its_point(ja):
    prob, model, p, x, v = _setup(ja, torch.float32)
    model.nll_and_grad(*p)
    with pytest.raises(RuntimeError, match="geovi_point"):
        model.geovi_apply("K", *v)
    model.geovi_point(*x)
    model.geovi_apply("K", *v)
    model.nll_and_grad(*p)  # moves the points: the operators need a new geovi_point
    with pytest.raises(RuntimeError, match="geovi_point"):
        model.geovi_apply("lh")
