"""GP sky, response, likelihood and the fused gradient / metric on the GPU
against the float64 oracle."""
import numpy as np
import pytest
import torch

from oracle import likelihood as olh
from oracle import model as om
from oracle import response as orsp
from oracle import sky as osky
from tests.problems import make_problem, powerlaw_amp, to_product

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


@pytest.mark.parametrize("N,H,E", [(64, 64, 1), (64, 48, 3), (256, 256, 4), (512, 512, 2),
                                   (1024, 1000, 2)])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_gp_forward_and_vjp(ja, N, H, E, dtype):
    rng = np.random.default_rng(N + H + E)
    amp = powerlaw_amp(N)
    amp_a = 0.5 * amp if E > 1 else None
    dnu = np.log(np.linspace(1, 3, E)) if E > 1 else None
    xr = rng.normal(size=(N, N))
    xa = rng.normal(size=(N, N)) if E > 1 else None
    want = osky.gp_sky(xr, xa, amp, amp_a, -2.0, -1.0, dnu if E > 1 else [0.0], crop=(H, H))
    sky = ja.CorrelatedFieldSky(N, (H, H), amp, amp_a, zm=-2.0, mu=-1.0, dnu=dnu, dtype=dtype)
    got = sky(torch.as_tensor(xr, device="cuda"),
              torch.as_tensor(xa, device="cuda") if xa is not None else None)
    tol = 1e-12 if dtype == torch.float64 else 2e-6
    assert _rel(got.double().cpu().numpy(), want) <= tol
    cot = rng.normal(size=want.shape)
    g_ref, g_alpha, _, _ = osky.gp_adjoint(want * cot, amp, amp_a, dnu if E > 1 else [0.0],
                                           grid_shape=(N, N))
    gr, ga = sky.vjp(torch.as_tensor(want, device="cuda"), torch.as_tensor(cot, device="cuda"))
    assert _rel(gr.double().cpu().numpy(), g_ref) <= tol
    if E > 1:
        assert _rel(ga.double().cpu().numpy(), g_alpha) <= tol


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_poisson_energy_and_gradient(ja, dtype):
    rng = np.random.default_rng(2)
    lam = rng.uniform(0.5, 50.0, size=10007)
    d = rng.poisson(lam).astype(np.float64)
    nll, g = ja.Poissonian(d).energy_and_gradient(lam.astype(np.float32) if dtype ==
                                                  torch.float32 else lam)
    want = olh.poisson_nll(lam, d)
    assert abs(nll - want) <= (1e-12 if dtype == torch.float64 else 1e-6) * abs(want)
    np.testing.assert_allclose(g, olh.poisson_grad(lam, d),
                               rtol=1e-12 if dtype == torch.float64 else 1e-5, atol=1e-6)


def test_exposure_matches_reference_fixture(ja):
    """jubik test_build_exposure_function.py fixtures (seeded numpy)."""
    exposures = np.random.default_rng(0).uniform(0.0, 3e3, size=3 * 100 ** 2).reshape(3, 100, 100)
    x = np.linspace(0.25, 1.75, num=100 * 100).reshape(100, 100)
    got = ja.build_exposure_function(exposures.copy(), 500)(x)
    want = orsp.build_exposure_function(exposures.copy(), 500)(x)
    np.testing.assert_array_equal(got, want)
    got = ja.build_exposure_function(exposures.copy(), None)(x)
    np.testing.assert_array_equal(got, exposures * x)


def test_exposure_sees_later_in_place_edits(ja):
    """The reference's closure multiplies by the caller's array as it is at
    each call (jubik/response.py:12-45): an in-place edit after the first call
    shows in the next one, for host arrays and device tensors alike."""
    rng = np.random.default_rng(3)
    x = rng.uniform(0.5, 1.5, size=(2, 16, 16))
    e = rng.uniform(1.0, 2.0, size=(2, 16, 16))
    f = ja.build_exposure_function(e)
    np.testing.assert_array_equal(f(x), e * x)
    e[0] *= 3.0
    np.testing.assert_array_equal(f(x), e * x)
    et = torch.as_tensor(e, device="cuda")
    ft = ja.build_exposure_function(et)
    xt = torch.as_tensor(x, device="cuda")
    assert torch.equal(ft(xt), et * xt)
    et[1] = 0.0
    assert torch.equal(ft(xt), et * xt)


def test_readout_matches_reference_fixture(ja):
    """jubik test_build_exposure_readout_function.py fixtures."""
    exposures = np.random.default_rng(0).uniform(0.0, 3e3, size=2 * 24 ** 2 * 3).reshape(
        2, 3, 24, 24)
    x = np.arange(2 * 3 * 24 * 24, dtype=float).reshape(2, 3, 24, 24) + 0.5
    keys = ('tm_1', 'tm_2')
    flags = exposures.copy()
    res = ja.build_readout_function(flags, 20, keys)(x)
    assert tuple(res.tree.keys()) == keys
    mask = exposures < 20
    for i, k in enumerate(keys):
        np.testing.assert_array_equal(res.tree[k], x[i][~mask[i]])
    # transpose scatters back
    ro = ja.build_readout_function(exposures.copy(), 20, keys)
    back = ro.transpose(res)
    np.testing.assert_array_equal(back, np.where(mask, 0.0, x))


@pytest.mark.parametrize("cfg", [
    dict(H=64, n=4, margin=8, E=1, M=1, seed=1),
    dict(H=64, n=4, margin=8, E=3, M=2, seed=2),
    dict(H=128, n=4, margin=16, E=2, M=1, seed=3, N=128),
    dict(H=512, n=4, margin=64, E=2, M=1, seed=4, N=512),  # L = 384, small-radix plans
])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_model_grad_matches_oracle(ja, cfg, dtype):
    prob, xi = make_problem(**cfg)
    want_nll, (want_r, want_a) = om.nll_and_grad(prob, *xi)
    model = to_product(ja, prob, dtype=dtype)
    nll, g_r, g_a = model.nll_and_grad(*[None if v is None else torch.as_tensor(v, device="cuda")
                                         for v in xi])
    got = g_r.double().cpu().numpy().ravel()
    want = want_r.ravel()
    if want_a is not None:
        got = np.concatenate([got, g_a.double().cpu().numpy().ravel()])
        want = np.concatenate([want, want_a.ravel()])
    err = _rel(got, want)
    nerr = abs(nll - want_nll) / abs(want_nll)
    print(f"{cfg} {dtype}: nll rel {nerr:.2e} grad rel {err:.2e}")
    if dtype == torch.float64:
        assert nerr <= 1e-12 and err <= 1e-10
    else:
        assert nerr <= 1e-6 and err <= 1e-5


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_model_metric_matches_oracle(ja, dtype):
    prob, xi = make_problem(H=64, n=4, margin=8, E=2, M=1, seed=5)
    rng = np.random.default_rng(6)
    v = (rng.normal(size=xi[0].shape), rng.normal(size=xi[1].shape))
    want_r, want_a = om.metric_product(prob, xi[0], xi[1], *v)
    model = to_product(ja, prob, dtype=dtype)
    model.nll_and_grad(*[torch.as_tensor(a, device="cuda") for a in xi])
    got_r, got_a = model.metric(*[torch.as_tensor(a, device="cuda") for a in v])
    got = np.concatenate([got_r.double().cpu().numpy().ravel(),
                          got_a.double().cpu().numpy().ravel()])
    want = np.concatenate([want_r.ravel(), want_a.ravel()])
    err = _rel(got, want)
    print(f"metric {dtype}: rel {err:.2e}")
    assert err <= (1e-10 if dtype == torch.float64 else 1e-5)
