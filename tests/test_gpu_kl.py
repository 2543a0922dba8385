"""KL minimisation on the device: the metric at the sample positions, the
device-resident CG (against dense solves built from the float64 oracle),
the Newton step, a short MGVI run, graph capture of the solver and of the
metric batches of several observations."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import model as om
from tests.problems import make_problem, to_product

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def _cat(a, b):
    return np.concatenate([np.ravel(a)] + ([] if b is None else [np.ravel(b)]))


def _positions(xi, S, seed, scale=0.3):
    """S antithetic sample positions around xi (pairs xi +- d)."""
    rng = np.random.default_rng(seed)
    d_r = rng.normal(size=(S // 2,) + xi[0].shape) * scale
    d_a = rng.normal(size=(S // 2,) + xi[1].shape) * scale if xi[1] is not None else None
    dr = np.stack([v for k in range(S // 2) for v in (d_r[k], -d_r[k])])
    da = None if d_a is None else np.stack([v for k in range(S // 2) for v in (d_a[k], -d_a[k])])
    return dr, da


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-10), (torch.float32, 1e-5)])
def test_metric_at_sample_positions(ja, dtype, tol):
    prob, xi = make_problem(H=64, n=4, margin=8, E=2, M=2, seed=61)
    S = 4
    model = to_product(ja, prob, dtype, max_samples=S)
    dr, da = _positions(xi, S, 1)
    pr, pa = xi[0][None] + dr, xi[1][None] + da
    model.nll_and_grad_batch(torch.as_tensor(pr, dtype=dtype, device="cuda"),
                             torch.as_tensor(pa, dtype=dtype, device="cuda"))
    rng = np.random.default_rng(2)
    v = (rng.normal(size=xi[0].shape), rng.normal(size=xi[1].shape))
    got_r, got_a = model.metric_samples(torch.as_tensor(v[0], dtype=dtype, device="cuda"),
                                        torch.as_tensor(v[1], dtype=dtype, device="cuda"))
    want_r = np.zeros(xi[0].shape)
    want_a = np.zeros(xi[1].shape)
    for s in range(S):
        mr, ma = om.metric_product(prob, pr[s], pa[s], v[0], v[1])
        want_r += mr
        want_a += ma
    assert _rel(got_r.double().cpu().numpy(), want_r) <= tol
    assert _rel(got_a.double().cpu().numpy(), want_a) <= tol


def _dense_kl_metric(prob, pr, pa):
    """I + 1/S sum_s J_s^T diag(1/lam_s) J_s over (xi_ref, xi_alpha)."""
    N = prob.amp_ref.shape[0]
    n = N * N
    S = pr.shape[0]
    M = np.eye(2 * n)
    for j in range(2 * n):
        v = np.zeros(2 * n)
        v[j] = 1.0
        for s in range(S):
            mr, ma = om.metric_product(prob, pr[s], pa[s], v[:n].reshape(N, N),
                                       v[n:].reshape(N, N))
            M[:, j] += np.concatenate([mr.ravel(), ma.ravel()]) / S
    return M


def test_kl_solve_and_newton_step_vs_dense(ja):
    prob, xi = make_problem(H=16, n=2, margin=4, E=2, M=1, seed=62)
    N, S = 16, 4
    model = to_product(ja, prob, torch.float64, max_samples=S)
    dr, da = _positions(xi, S, 3)
    pr, pa = xi[0][None] + dr, xi[1][None] + da
    dev = lambda a: torch.as_tensor(a, dtype=torch.float64, device="cuda")
    value, g, d, info = ja.kl.newton_step(model, dev(xi[0]), dev(xi[1]), dev(dr), dev(da),
                                          cg_rtol=1e-12, cg_maxiter=2000)
    # the KL value and gradient against the oracle
    want_val = 0.0
    want_g = np.zeros(2 * N * N)
    for s in range(S):
        nll, (gr, ga) = om.nll_and_grad(prob, pr[s], pa[s])
        want_val += (nll + 0.5 * (np.sum(pr[s] ** 2) + np.sum(pa[s] ** 2))) / S
        want_g += (_cat(gr, ga) + _cat(pr[s], pa[s])) / S
    assert abs(value - want_val) <= 1e-10 * abs(want_val)
    assert _rel(_cat(g[0].cpu().numpy(), g[1].cpu().numpy()), want_g) <= 1e-10
    # the Newton direction solves the dense KL metric system
    Md = _dense_kl_metric(prob, pr, pa)
    want_d = np.linalg.solve(Md, -want_g)
    got_d = _cat(d[0].cpu().numpy(), d[1].cpu().numpy())
    assert _rel(got_d, want_d) <= 1e-8
    assert info["rel_residual"] <= 1e-12


def test_newton_cg_decreases_the_kl(ja):
    prob, xi = make_problem(H=32, n=2, margin=8, E=2, M=1, seed=63, expo_range=(50.0, 50.0))
    S = 4
    model = to_product(ja, prob, torch.float64, max_samples=S)
    dr, da = _positions(xi, S, 4, scale=0.2)
    dev = lambda a: torch.as_tensor(a, dtype=torch.float64, device="cuda")
    start = dev(xi[0]) + 0.5 * dev(np.random.default_rng(5).normal(size=xi[0].shape))
    xr, xa, hist = ja.newton_cg(model, start, dev(xi[1]), dev(dr), dev(da), xtol=1e-10,
                                maxiter=6, cg_rtol=1e-8, cg_maxiter=500)
    kl = hist["kl"]
    assert len(kl) >= 3
    assert all(b < a for a, b in zip(kl, kl[1:])), kl
    # Newton converges: the last accepted steps are full steps and the KL
    # change shrinks
    assert hist["step"][-1] == 1.0
    assert abs(kl[-1] - kl[-2]) < abs(kl[1] - kl[0])


def test_optimize_kl_short_run(ja):
    prob, xi = make_problem(H=32, n=2, margin=8, E=2, M=1, seed=64, expo_range=(50.0, 50.0))
    model = to_product(ja, prob, torch.float32, max_samples=4)
    x0 = np.zeros_like(xi[0]), np.zeros_like(xi[1])
    xr, xa, sr, sa, hist = ja.optimize_kl(model, x0[0], x0[1], n_samples=4, n_iterations=2,
                                          seed=1, maxiter=4, draw_rtol=1e-4, draw_maxiter=400)
    assert sr.shape == (4, 32, 32)
    # each global iteration lowers its sampled KL, and the posterior mean
    # moves towards the truth's neighbourhood (the data's position)
    for h in hist:
        assert h["kl"][-1] < h["kl"][0]
    err0 = np.linalg.norm(x0[0] - xi[0])
    err1 = np.linalg.norm(xr.double().cpu().numpy() - xi[0])
    assert err1 < err0


def test_solver_is_capturable(ja):
    """jubik_model_metric_solve with check_every = 0 launches a fixed kernel
    sequence without host synchronisation: captured in a graph and replayed
    (with eager work in between) it gives the eager result bit for bit."""
    from jubik_amd import _lib
    prob, xi = make_problem(H=32, n=2, margin=8, E=2, M=1, seed=65)
    S, N = 2, 32
    model = to_product(ja, prob, torch.float32, max_samples=S).reserve_solver()
    model.nll_and_grad(xi[0], xi[1])
    rng = np.random.default_rng(6)
    b_r = torch.as_tensor(rng.normal(size=(S, N, N)), dtype=torch.float32, device="cuda")
    b_a = torch.as_tensor(rng.normal(size=(S, N, N)), dtype=torch.float32, device="cuda")
    x_r, x_a = torch.empty_like(b_r), torch.empty_like(b_a)
    info = torch.zeros(2 * S, dtype=torch.float64, device="cuda")
    sky = model.sky
    ptr = lambda t: ctypes.c_void_p(t.data_ptr())

    def solve():
        _lib.call("jubik_model_metric_solve", model._model, S, ptr(sky.amp_ref),
                  ptr(sky.amp_alpha), ptr(sky.dnu_dev), ptr(b_r), ptr(b_a), ptr(x_r), ptr(x_a),
                  25, 1e-9, 0.0, 0, ptr(info),
                  ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        solve()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    want = (x_r.clone(), x_a.clone(), info.clone())
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        solve()
    for k in range(3):
        x_r.zero_()
        model.metric_batch(b_r, b_a)  # eager work between replays
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(x_r, want[0]) and torch.equal(x_a, want[1])
        assert torch.equal(info, want[2])


def test_metric_graph_of_uneven_grids(ja):
    """The antithetic metric products of three observations on uneven grids
    (128^2, 256^2, 512^2) in one graph equal eager calls."""
    geoms = [dict(H=128, n=4, margin=16, E=1, M=1, seed=66),
             dict(H=256, n=4, margin=32, E=1, M=1, seed=67),
             dict(H=512, n=8, margin=32, E=1, M=1, seed=68)]
    S = 4
    models, xis = [], []
    for g in geoms:
        prob, xi = make_problem(**g)
        m = to_product(ja, prob, torch.float32, max_samples=S)
        m.nll_and_grad(xi[0], xi[1])
        models.append(m)
        xis.append(xi)
    graph = ja.capture_metric_batch([(m, S) for m in models])
    rng = np.random.default_rng(9)
    for _ in range(2):
        vs = []
        for sl, m in zip(graph.slots, models):
            v = torch.as_tensor(rng.normal(size=(S // 2, m.sky.N, m.sky.N)), dtype=torch.float32,
                                device="cuda")
            v = torch.cat([v, -v])
            sl.v_ref.copy_(v)
            vs.append(v)
        graph.launch()
        torch.cuda.synchronize()
        for sl, m, v in zip(graph.slots, models, vs):
            want, _ = m.metric_batch(v)
            torch.cuda.synchronize()
            assert torch.equal(sl.out_ref, want)
            # antithetic pairs give opposite products
            assert torch.allclose(sl.out_ref[: S // 2], -sl.out_ref[S // 2:], rtol=1e-6,
                                  atol=1e-6 * float(sl.out_ref.abs().max()))
