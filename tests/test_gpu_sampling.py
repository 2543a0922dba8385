"""Linear (MGVI) sample draws: the metric's square root, the CG solve and
the antithetic draw, against the float64 oracle and a dense solve."""
import numpy as np
import pytest
import torch

from oracle import model as om
from tests.problems import make_problem, to_product

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def test_sqrt_metric_vs_oracle(ja):
    """J^T diag(lam^-1/2) eta equals the oracle's adjoint of eta / sqrt(lam)."""
    prob, xi = make_problem(H=64, n=4, margin=8, E=2, M=2, seed=51)
    S = 2
    model = to_product(ja, prob, torch.float64, max_samples=S)
    model.nll_and_grad(xi[0], xi[1])
    rng = np.random.default_rng(1)
    noise = rng.normal(size=(S, 2, 2, 64, 64))
    got_r, got_a = model.sqrt_metric_batch(noise)
    s = om.sky(prob, xi[0], xi[1])
    lam = om.response(prob, s)
    for k in range(S):
        cot = []
        for mi, (mod, lm) in enumerate(zip(prob.modules, lam)):
            c = np.zeros(lm.shape)
            c[mod.keep] = noise[k, mi][mod.keep] / np.sqrt(lm[mod.keep])
            cot.append(c)
        want_r, want_a = om._backward(prob, s, cot)
        assert _rel(got_r[k], want_r) <= 1e-11
        assert _rel(got_a[k], want_a) <= 1e-11


def _dense_metric(prob, xi):
    """I + J^T diag(1/lam) J as a dense matrix over (xi_ref, xi_alpha)."""
    N = prob.amp_ref.shape[0]
    n = N * N
    M = np.eye(2 * n)
    for j in range(2 * n):
        v = np.zeros(2 * n)
        v[j] = 1.0
        mr, ma = om.metric_product(prob, xi[0], xi[1], v[:n].reshape(N, N), v[n:].reshape(N, N))
        M[:, j] += np.concatenate([mr.ravel(), ma.ravel()])
    return M


def test_metric_solve_vs_dense(ja):
    prob, xi = make_problem(H=16, n=2, margin=4, E=2, M=1, seed=52)
    N = 16
    model = to_product(ja, prob, torch.float64, max_samples=2)
    model.nll_and_grad(xi[0], xi[1])
    rng = np.random.default_rng(2)
    b = rng.normal(size=(2, 2 * N * N))
    (xr, xa), info = ja.metric_solve(model, b[:, :N * N].reshape(2, N, N),
                                     b[:, N * N:].reshape(2, N, N), rtol=1e-11, maxiter=2000)
    Md = _dense_metric(prob, xi)
    assert np.allclose(Md, Md.T, rtol=1e-10, atol=1e-12)
    for k in range(2):
        want = np.linalg.solve(Md, b[k])
        got = np.concatenate([xr[k].cpu().numpy().ravel(), xa[k].cpu().numpy().ravel()])
        assert _rel(got, want) <= 1e-8
    assert np.all(info["rel_residual"] <= 1e-11)


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-10), (torch.float32, 1e-4)])
def test_draw_linear_samples(ja, dtype, tol):
    # a small, moderately conditioned problem so the CG reaches tol
    prob, xi = make_problem(H=32, n=2, margin=8, E=2, M=1, seed=53, expo_range=(20.0, 20.0))
    model = to_product(ja, prob, dtype, max_samples=3)
    sr, sa, info = ja.draw_linear_samples(model, xi[0], xi[1], n_samples=6, seed=7,
                                          rtol=tol, maxiter=3000)
    assert sr.shape == (6, 32, 32) and sa.shape == (6, 32, 32)
    assert np.all(info["rel_residual"] <= tol)
    xr = torch.as_tensor(xi[0], device="cuda").to(dtype)
    xa = torch.as_tensor(xi[1], device="cuda").to(dtype)
    scale = float(xr.abs().max())
    for k in range(3):
        # antithetic pairs mirror the base point
        assert float((sr[2 * k] + sr[2 * k + 1] - 2 * xr).abs().max()) <= 4 * tol * scale + 1e-6
        assert float((sa[2 * k] + sa[2 * k + 1] - 2 * xa).abs().max()) <= 4 * tol * scale + 1e-6
    # the same seed reproduces the draw bit for bit, another seed does not
    sr2, _, _ = ja.draw_linear_samples(model, xi[0], xi[1], n_samples=6, seed=7, rtol=tol,
                                       maxiter=3000)
    assert torch.equal(sr, sr2)
    sr3, _, _ = ja.draw_linear_samples(model, xi[0], xi[1], n_samples=6, seed=8, rtol=tol,
                                       maxiter=3000)
    assert not torch.equal(sr, sr3)


def test_conjugate_gradient_on_a_known_matrix(ja):
    """The batched CG recurrences on an explicit SPD matrix (device tensors)."""
    rng = np.random.default_rng(3)
    A = rng.normal(size=(40, 40))
    A = A @ A.T + 40 * np.eye(40)
    At = torch.as_tensor(A, device="cuda")
    b = rng.normal(size=(3, 40))
    x, info = ja.conjugate_gradient(lambda v: (v[0] @ At.T,), (torch.as_tensor(b, device="cuda"),),
                                    rtol=1e-12, maxiter=200)
    want = np.linalg.solve(A, b.T).T
    assert _rel(x[0].cpu().numpy(), want) <= 1e-10
