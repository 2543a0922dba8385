"""Parts of the model's surface beyond one plain gradient: a pixel area a != 1
through every path that carries it, the lam <= 0 domain check, the Poisson
likelihood's metric and its left square root, and the safety of captured
graphs and metric base points when the sky parameters change."""
import numpy as np
import pytest
import torch

from oracle import convolve as ocv
from oracle import model as om
from oracle.data import Domain
from tests.problems import make_problem, psf_windows, to_product

pytestmark = pytest.mark.gpu

AREA = 0.37  # (fov / sdim)^2 of chandra_response.py:136, erosita_response.py:307


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def _dev(a, dtype=None):
    return None if a is None else torch.as_tensor(a, dtype=dtype, device="cuda")


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-10), (torch.float32, 1e-5)])
def test_pixel_area_gradient_and_metric(ja, dtype, tol):
    prob, xi = make_problem(H=64, n=4, margin=8, E=3, M=2, seed=31, pixel_area=AREA)
    want_nll, (wr, wa) = om.nll_and_grad(prob, *xi)
    model = to_product(ja, prob, dtype)
    nll, gr, ga = model.nll_and_grad(_dev(xi[0], dtype), _dev(xi[1], dtype))
    assert abs(nll - want_nll) <= tol * abs(want_nll)
    assert _rel(gr.double().cpu().numpy(), wr) <= tol
    assert _rel(ga.double().cpu().numpy(), wa) <= tol
    rng = np.random.default_rng(3)
    vr, va = rng.normal(size=xi[0].shape), rng.normal(size=xi[1].shape)
    mr, ma = om.metric_product(prob, xi[0], xi[1], vr, va)
    got_r, got_a = model.metric(_dev(vr, dtype), _dev(va, dtype))
    assert _rel(got_r.double().cpu().numpy(), mr) <= tol
    assert _rel(got_a.double().cpu().numpy(), ma) <= tol


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-10), (torch.float32, 1e-5)])
def test_pixel_area_hyper_and_batch(ja, dtype, tol):
    prob, xi = make_problem(H=64, n=4, margin=8, E=2, M=1, seed=32, N=128, pixel_area=AREA)
    want = om.nll_and_grad_hyper(prob, *xi)
    model = to_product(ja, prob, dtype, max_samples=2)
    got = model.nll_and_grad_hyper(_dev(xi[0], dtype), _dev(xi[1], dtype))
    assert abs(got[0] - want[0]) <= tol * abs(want[0])
    for g, w in zip(got[1:5], want[1:5]):
        assert _rel(g.double().cpu().numpy(), w) <= tol
    for g, w in zip(got[5:], want[5:]):
        assert abs(g - w) <= max(tol * 10, 1e-8) * abs(w)
    # two samples in one pass: the second at a shifted position
    x2 = (xi[0] + 0.1, xi[1] - 0.1)
    nll_b, gr_b, ga_b = model.nll_and_grad_batch(
        torch.as_tensor(np.stack([xi[0], x2[0]]), dtype=dtype, device="cuda"),
        torch.as_tensor(np.stack([xi[1], x2[1]]), dtype=dtype, device="cuda"))
    w2 = om.nll_and_grad(prob, *x2)
    assert abs(nll_b[1] - w2[0]) <= tol * abs(w2[0])
    assert _rel(gr_b[1].double().cpu().numpy(), w2[1][0]) <= tol


def test_pixel_area_graph_matches_eager(ja):
    prob, xi = make_problem(H=64, n=4, margin=8, E=2, M=2, seed=33, pixel_area=AREA)
    model = to_product(ja, prob, torch.float32)
    graph = model.capture_grad(0)
    graph.xi_ref.copy_(_dev(xi[0]))
    graph.xi_alpha.copy_(_dev(xi[1]))
    for _ in range(2):
        graph.launch()
        torch.cuda.synchronize()
        got = (float(graph.nll.item()), graph.g_ref.cpu().numpy().copy())
        nll, gr, _ = model.nll_and_grad(_dev(xi[0], torch.float32), _dev(xi[1], torch.float32))
        assert got[0] == nll and np.array_equal(got[1], gr.cpu().numpy())
    want_nll, (wr, _) = om.nll_and_grad(prob, *xi)
    assert _rel(got[1], wr) <= 1e-5


def test_lambda_domain_check(ja):
    """Counts read out where the exposure is zero give lam = 0: the NLL is
    inf there (as the reference's), the evaluation counts those pixels and
    the checked calls raise FloatingPointError."""
    prob, xi = make_problem(H=64, n=4, margin=8, E=2, M=1, seed=34)
    mod = prob.modules[0]
    holes = np.zeros(mod.keep.shape, dtype=bool)
    holes[0, 10, 3:8] = True
    holes[1, 40, 20:23] = True
    expo = mod.exposure.copy()
    expo[holes] = 0.0
    counts = np.where(mod.keep | holes, 1, -1).astype(np.int32)
    counts[mod.keep & ~holes] = 3
    model = to_product(ja, prob, torch.float64)
    bad = ja.XrayModel(model.sky, model.psf, prob.pixel_area, expo[None], counts[None])
    with pytest.raises(FloatingPointError, match="lambda <= 0"):
        bad.nll_and_grad(_dev(xi[0]), _dev(xi[1]))
    nll, _, _ = bad.nll_and_grad_device(_dev(xi[0]), _dev(xi[1]))
    with pytest.raises(FloatingPointError):
        bad.check_domain()
    assert list(bad.check_domain(raise_error=False))[0] == int(holes.sum())
    assert not np.isfinite(float(nll.item()))
    # the masked problem passes and reports zero such pixels
    model.nll_and_grad(_dev(xi[0]), _dev(xi[1]))
    assert list(model.check_domain())[0] == 0


def test_lambda_domain_counts_per_sample(ja):
    prob, xi = make_problem(H=64, n=4, margin=8, E=1, M=1, seed=35)
    mod = prob.modules[0]
    expo = mod.exposure.copy()
    expo[0, 5, 5:9] = 0.0  # four read-out pixels without exposure
    counts = np.where(mod.keep, 2, -1).astype(np.int32)
    counts[0, 5, 5:9] = 2
    model = to_product(ja, prob, torch.float32, max_samples=3)
    bad = ja.XrayModel(model.sky, model.psf, prob.pixel_area, expo[None], counts[None],
                       max_samples=3)
    xs = torch.as_tensor(np.stack([xi[0]] * 3), dtype=torch.float32, device="cuda")
    with pytest.raises(FloatingPointError, match="4 at sample 0, 4 at sample 1, 4 at sample 2"):
        bad.nll_and_grad_batch(xs)


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-13), (torch.float32, 2e-7)])
def test_poissonian_metric_and_left_sqrt(ja, dtype, tol):
    rng = np.random.default_rng(36)
    lam = rng.uniform(0.3, 40.0, size=5003)
    t = rng.normal(size=5003)
    d = rng.poisson(lam).astype(np.float64)
    lh = ja.Poissonian(d)
    cast = (lambda a: a.astype(np.float32)) if dtype == torch.float32 else (lambda a: a)
    np.testing.assert_allclose(lh.metric(cast(lam), cast(t)), t / lam, rtol=tol * 4)
    np.testing.assert_allclose(lh.left_sqrt_metric(cast(lam), cast(t)), t / np.sqrt(lam),
                               rtol=tol * 4)
    # Vector-valued data (one key per module) keeps its structure
    v = ja.Vector({"tm_1": lam[:100], "tm_2": lam[100:300]})
    tv = ja.Vector({"tm_1": t[:100], "tm_2": t[100:300]})
    out = ja.Poissonian(ja.Vector({"tm_1": d[:100], "tm_2": d[100:300]})).metric(v, tv)
    np.testing.assert_allclose(out.tree["tm_2"], t[100:300] / lam[100:300], rtol=1e-13)


def test_poissonian_metric_through_linpatch(ja):
    """Amended with a linear PSF operator the metric is A^T diag(1/Ax) A t
    and the left square root A^T diag((Ax)^-1/2) t (against the oracle's
    linpatch and its adjoint)."""
    rng = np.random.default_rng(37)
    H, n, m = 64, 4, 8
    win = psf_windows(rng, n, 1, 2 * (H // n))[:, 0]
    op = ja.LinpatchOperator((H, H), n, m, dtype=torch.float64, kernel_per_energy=False)
    op.set_kernel(win, layout="window")
    dom = Domain((H, H), (1.0, 1.0))
    x = rng.uniform(1.0, 2.0, size=(H, H))
    t = rng.normal(size=(H, H))
    d = rng.poisson(10.0, size=(H, H)).astype(np.float64)
    lh = ja.Poissonian(d).amend(op)
    ax = ocv.linpatch_convolve_window(x, dom, win, n, m)
    at = ocv.linpatch_convolve_window(t, dom, win, n, m)
    want = ocv.linpatch_adjoint_window(at / ax, dom, win, n, m)
    got = lh.metric(x, t)
    assert _rel(got, want) <= 1e-12
    want_s = ocv.linpatch_adjoint_window(t / np.sqrt(ax), dom, win, n, m)
    assert _rel(lh.left_sqrt_metric(x, t), want_s) <= 1e-12


def test_set_parameters_keeps_graphs_and_base_point_safe(ja):
    prob, xi = make_problem(H=64, n=4, margin=8, E=2, M=1, seed=38)
    model = to_product(ja, prob, torch.float32)
    sky = model.sky
    graph = model.capture_grad(0)
    graph.xi_ref.copy_(_dev(xi[0]))
    graph.xi_alpha.copy_(_dev(xi[1]))
    graph.launch()
    ptr = sky.amp_ref.data_ptr()
    sky.set_parameters(amp_ref=1.3 * prob.amp_ref, amp_alpha=0.7 * prob.amp_alpha)
    assert sky.amp_ref.data_ptr() == ptr  # overwritten in place
    with pytest.raises(RuntimeError, match="sky parameters changed"):
        model.metric(_dev(xi[0], torch.float32), _dev(xi[1], torch.float32))
    graph.launch()  # reads the new amplitude fields
    torch.cuda.synchronize()
    got = (float(graph.nll.item()), graph.g_ref.cpu().numpy().copy())
    nll, gr, _ = model.nll_and_grad(_dev(xi[0], torch.float32), _dev(xi[1], torch.float32))
    assert got[0] == nll and np.array_equal(got[1], gr.cpu().numpy())
    model.metric(_dev(xi[0], torch.float32), _dev(xi[1], torch.float32))  # base point current
    sky.set_parameters(zm=-1.5)
    with pytest.raises(RuntimeError, match="capture it again"):
        graph.launch()
    # a fresh capture sees the new offset and matches the oracle
    g2 = model.capture_grad(0)
    g2.xi_ref.copy_(_dev(xi[0]))
    g2.xi_alpha.copy_(_dev(xi[1]))
    g2.launch()
    torch.cuda.synchronize()
    from dataclasses import replace
    p2 = replace(prob, amp_ref=1.3 * prob.amp_ref, amp_alpha=0.7 * prob.amp_alpha, zm=-1.5)
    want_nll, (wr, _) = om.nll_and_grad(p2, *xi)
    assert abs(float(g2.nll.item()) - want_nll) <= 1e-5 * abs(want_nll)
    assert _rel(g2.g_ref.double().cpu().numpy(), wr) <= 1e-5
