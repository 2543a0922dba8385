"""Pins of the oracle that the reference's tests leave open.

linpatch_convolve: an independent loop version (oracle/brute.py), the
delta-PSF coverage identity, a single shared PSF against scipy's linear
convolution, linearity and the hand adjoint.  GP: Hartley identities and the
amplitude normalisation.  Likelihood and model: closed forms and central
finite differences in float64.
"""
import numpy as np
import pytest
import scipy.signal

from oracle import brute
from oracle import convolve as cv
from oracle import likelihood as lh
from oracle import model as om
from oracle import sky as gp
from oracle.data import Domain
from tests.problems import make_problem, powerlaw_amp, psf_windows

GEOMS = [(16, 4, 4), (16, 4, 2), (18, 4, 3), (20, 3, 2), (14, 4, 2), (24, 3, 5)]


@pytest.mark.parametrize("H,n,m", GEOMS)
def test_linpatch_matches_loop_version(H, n, m):
    rng = np.random.default_rng(H * 10 + n + m)
    x = rng.normal(size=(H, H))
    win = rng.random((n * n, 2 * (H // n), 2 * (H // n)))
    got = cv.linpatch_convolve_window(x, Domain((H, H), (1.0, 1.0)), win, n, m)
    want = np.array(brute.linpatch_loops(x.tolist(), win.tolist(), n, m))
    np.testing.assert_allclose(got, want, atol=1e-12)


@pytest.mark.parametrize("H,n,m", [(16, 4, 4), (18, 4, 3)])
def test_full_frame_kernel_equals_window_form(H, n, m):
    rng = np.random.default_rng(5)
    x = rng.normal(size=(2, H, H))
    win = rng.random((n * n, 2, 2 * (H // n), 2 * (H // n)))
    full = cv.window_to_full_frame(win, (H, H))
    np.testing.assert_array_equal(cv.cut_kernel(full, (2, H, H), n), win)
    dom = Domain((2, H, H), (0.3, 0.3, 0.3))
    np.testing.assert_allclose(cv.linpatch_convolve(x, dom, full, n, m),
                               cv.linpatch_convolve_window(x, dom, win, n, m), atol=1e-13)


def _coverage_1d(N, n):
    """sum over windows of the tent weight at each pixel, by explicit loops."""
    d = N // n
    h = d // 2
    w = 2 * d
    ramp = [t / (d - 1) for t in range(d)] if d > 1 else [0.0]
    tent = ramp + ramp[::-1]
    lim = N + 2 * h - w
    cov = np.zeros(N)
    for i in range(n):
        s = min(i * d, lim)
        for t in range(w):
            j = s + t - h
            if 0 <= j < N:
                cov[j] += tent[t]
    return cov


@pytest.mark.parametrize("H,n,m", [(32, 4, 4), (30, 4, 3), (64, 8, 8)])
def test_delta_psf_gives_coverage_times_input(H, n, m):
    """A delta PSF in every patch leaves each window unchanged, so the output
    is the tent coverage map times x; the coverage is one wherever two
    windows overlap along both axes (partition of unity)."""
    rng = np.random.default_rng(H)
    x = rng.normal(size=(H, H))
    w = 2 * (H // n)
    win = np.zeros((n * n, w, w))
    win[:, w // 2, w // 2] = 1.0
    got = cv.linpatch_convolve_window(x, Domain((H, H), (1.0, 1.0)), win, n, m)
    c = _coverage_1d(H, n)
    np.testing.assert_allclose(got, np.outer(c, c) * x, atol=1e-12)
    d = H // n
    inner = slice(d // 2 + d, H - d // 2 - d)
    np.testing.assert_allclose(np.outer(c, c)[inner, inner], 1.0, atol=1e-12)


def test_shared_small_psf_is_a_linear_convolution():
    """One PSF for all patches with support within the margin: no wrap-around,
    so the result is PSF * (coverage x) as a linear 'same' convolution."""
    H, n, m = 48, 4, 8
    rng = np.random.default_rng(11)
    x = rng.normal(size=(H, H))
    w = 2 * (H // n)
    k = rng.random((7, 7))
    k /= k.sum()
    win = np.zeros((n * n, w, w))
    win[:, w // 2 - 3:w // 2 + 4, w // 2 - 3:w // 2 + 4] = k
    got = cv.linpatch_convolve_window(x, Domain((H, H), (1.0, 1.0)), win, n, m)
    c = _coverage_1d(H, n)
    want = scipy.signal.convolve2d(np.outer(c, c) * x, k, mode="same")
    np.testing.assert_allclose(got, want, atol=1e-12)


def test_linpatch_linearity_and_adjoint():
    H, n, m, E = 40, 4, 6, 2
    rng = np.random.default_rng(4)
    dom = Domain((E, H, H), (1.0, 1.0, 1.0))
    win = psf_windows(rng, n, E, 2 * (H // n))
    a, b = rng.normal(size=(E, H, H)), rng.normal(size=(E, H, H))
    f = lambda v: cv.linpatch_convolve_window(v, dom, win, n, m)
    np.testing.assert_allclose(f(2 * a + b), 2 * f(a) + f(b), atol=1e-12)
    g = rng.normal(size=(E, H, H))
    lhs = np.vdot(g, f(a))
    rhs = np.vdot(cv.linpatch_adjoint_window(g, dom, win, n, m), a)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_linpatch_kernel_normalisation_cancels_dvol():
    """The unit-sum normalisation divides by sum * dvol and convolve
    multiplies by dvol: the pixel distances drop out."""
    H, n, m = 16, 4, 4
    rng = np.random.default_rng(9)
    x = rng.normal(size=(H, H))
    win = rng.random((n * n, 8, 8))
    a = cv.linpatch_convolve_window(x, Domain((H, H), (1.0, 1.0)), win, n, m)
    b = cv.linpatch_convolve_window(x, Domain((H, H), (0.37, 2.5)), 3.0 * win, n, m)
    np.testing.assert_allclose(a, b, atol=1e-12)


# ---------------------------------------------------------------- GP
def test_hartley_identities():
    N = 16
    rng = np.random.default_rng(0)
    delta = np.zeros((N, N))
    delta[0, 0] = 1.0
    np.testing.assert_allclose(gp.hartley(delta), 1.0)
    u = rng.normal(size=(N, N))
    np.testing.assert_allclose(gp.hartley(gp.hartley(u)), N * N * u, atol=1e-10)
    v = rng.normal(size=(N, N))
    assert abs(np.vdot(v, gp.hartley(u)) - np.vdot(gp.hartley(v), u)) < 1e-9
    np.testing.assert_allclose(np.sum(gp.hartley(u) ** 2), N * N * np.sum(u ** 2))


def test_hartley_is_cas_transform():
    """HT(u)[x] = sum_k u[k] (cos - sin)(2 pi k.x / N), by direct summation."""
    N = 6
    u = np.random.default_rng(1).normal(size=(N, N))
    k = np.arange(N)
    th = 2 * np.pi * (k[:, None, None, None] * k[None, None, :, None] +
                      k[None, :, None, None] * k[None, None, None, :]) / N
    want = np.einsum("ab,abxy->xy", u, np.cos(th) - np.sin(th))
    np.testing.assert_allclose(gp.hartley(u), want, atol=1e-12)


def test_powerlaw_amplitude_pixel_variance():
    """Var of HT(amp xi) per pixel averages to sum(amp^2) = fluct^2."""
    N, f = 32, 1.0
    amp = powerlaw_amp(N, fluct=f)
    assert amp[0, 0] == 0.0
    assert np.sum(amp ** 2) == pytest.approx(f ** 2)
    rng = np.random.default_rng(2)
    samples = np.stack([gp.hartley(amp * rng.normal(size=(N, N))) for _ in range(400)])
    assert np.var(samples) == pytest.approx(f ** 2, rel=0.08)


def test_spectral_composition_closed_forms():
    """The spectral sky's energy dependence in closed form
    (spectral_product_mf_sky.py:206-219): at xi = 0 every pixel holds
    log s_e = zm + mu dnu_e; with only xi_ref set the energies differ by
    mu (dnu_e - dnu_0) everywhere (the reference field carries no dnu);
    with only xi_alpha set log s_e - zm = (HT(amp_a xi_a) + mu) dnu_e, which
    a direct cas-sum gives for a single excited mode."""
    N = 8
    zm, mu = -2.0, -1.3
    dnu = np.log([1.0, 1.6, 2.5, 4.0])
    amp = powerlaw_amp(N)
    zero = np.zeros((N, N))
    ls = gp.gp_log_sky(zero, zero, amp, 0.5 * amp, zm, mu, dnu)
    np.testing.assert_allclose(ls, np.broadcast_to((zm + mu * dnu)[:, None, None], ls.shape),
                               rtol=0, atol=1e-15)
    rng = np.random.default_rng(4)
    xr = rng.normal(size=(N, N))
    ls = gp.gp_log_sky(xr, zero, amp, 0.5 * amp, zm, mu, dnu)
    diff = ls - ls[0]
    np.testing.assert_allclose(diff, np.broadcast_to((mu * (dnu - dnu[0]))[:, None, None],
                                                     diff.shape), atol=1e-13)
    # one excited mode k = (1, 2) of the spectral-index field
    xa = np.zeros((N, N))
    xa[1, 2] = 1.5
    x = np.arange(N)
    cas = (np.cos(2 * np.pi * (x[:, None] * 1 + x[None, :] * 2) / N) -
           np.sin(2 * np.pi * (x[:, None] * 1 + x[None, :] * 2) / N))
    t_alpha = 0.5 * amp[1, 2] * 1.5 * cas
    ls = gp.gp_log_sky(zero, xa, amp, 0.5 * amp, zm, mu, dnu)
    np.testing.assert_allclose(ls - zm, (t_alpha[None] + mu) * dnu[:, None, None], atol=1e-13)


def test_gp_adjoint_dot_product():
    N, E = 16, 3
    rng = np.random.default_rng(7)
    amp = powerlaw_amp(N)
    dnu = np.log([1.0, 1.7, 2.9])
    xr, xa = rng.normal(size=(N, N)), rng.normal(size=(N, N))
    log_s = gp.gp_log_sky(xr, xa, amp, 0.5 * amp, 0.0, 0.0, dnu)
    u = rng.normal(size=log_s.shape)
    gr, ga, _, _ = gp.gp_adjoint(u, amp, 0.5 * amp, dnu)
    assert abs(np.vdot(u, log_s) - (np.vdot(gr, xr) + np.vdot(ga, xa))) < 1e-9 * abs(
        np.vdot(u, log_s))


# ---------------------------------------------------------------- likelihood
def test_poisson_closed_form_and_stationary_point():
    lam = np.array([1.0, 2.0, 4.0])
    d = np.array([0.0, 1.0, 3.0])
    want = 7.0 - (np.log(2.0) + 3 * np.log(4.0))
    assert lh.poisson_nll(lam, d) == pytest.approx(want, rel=1e-15)
    np.testing.assert_allclose(lh.poisson_grad(d + 0.0 + (d == 0), d)[1:], 0.0)


def test_poisson_gradient_finite_differences():
    rng = np.random.default_rng(3)
    lam = rng.uniform(0.5, 5, 20)
    d = rng.poisson(lam).astype(float)
    g = lh.poisson_grad(lam, d)
    h = 1e-6
    for i in range(0, 20, 5):
        e = np.zeros(20)
        e[i] = h
        fd = (lh.poisson_nll(lam + e, d) - lh.poisson_nll(lam - e, d)) / (2 * h)
        assert fd == pytest.approx(g[i], rel=1e-7, abs=1e-8)


# ---------------------------------------------------------------- model
@pytest.mark.parametrize("cfg", [dict(H=32, n=4, margin=4, E=1, M=1, seed=1),
                                 dict(H=32, n=4, margin=4, E=3, M=2, seed=2),
                                 dict(H=32, n=4, margin=4, E=2, M=2, seed=3, pixel_area=0.37)])
def test_model_gradient_finite_differences(cfg):
    prob, xi = make_problem(**cfg)
    _, (gr, ga) = om.nll_and_grad(prob, *xi)
    rng = np.random.default_rng(0)
    v = (rng.normal(size=xi[0].shape), None if xi[1] is None else rng.normal(size=xi[1].shape))
    h = 1e-6
    plus = (xi[0] + h * v[0], None if xi[1] is None else xi[1] + h * v[1])
    minus = (xi[0] - h * v[0], None if xi[1] is None else xi[1] - h * v[1])
    fd = (om.nll(prob, *plus) - om.nll(prob, *minus)) / (2 * h)
    an = np.vdot(gr, v[0]) + (0.0 if ga is None else np.vdot(ga, v[1]))
    assert fd == pytest.approx(an, rel=1e-6)


@pytest.mark.parametrize("cfg", [dict(H=32, n=4, margin=4, E=1, M=1, seed=5, N=40),
                                 dict(H=32, n=4, margin=4, E=3, M=2, seed=6),
                                 dict(H=32, n=4, margin=4, E=2, M=1, seed=7, pixel_area=0.37)])
def test_hyper_gradients_finite_differences(cfg):
    """Amplitude-field and offset gradients against central differences of
    the NLL in the amplitude fields, zm and mu_alpha."""
    from dataclasses import replace
    prob, xi = make_problem(**cfg)
    nll, gr, ga, gar, gaa, gzm, gmu = om.nll_and_grad_hyper(prob, *xi)
    want_nll, (wr, wa) = om.nll_and_grad(prob, *xi)
    assert nll == want_nll
    np.testing.assert_array_equal(gr, wr)
    h = 1e-6
    f = lambda **kw: om.nll(replace(prob, **kw), *xi)
    rng = np.random.default_rng(2)
    v = rng.normal(size=prob.amp_ref.shape)
    fd = (f(amp_ref=prob.amp_ref + h * v) - f(amp_ref=prob.amp_ref - h * v)) / (2 * h)
    assert fd == pytest.approx(np.vdot(gar, v), rel=1e-6)
    fd = (f(zm=prob.zm + h) - f(zm=prob.zm - h)) / (2 * h)
    assert fd == pytest.approx(gzm, rel=1e-6)
    if prob.amp_alpha is None:
        assert ga is None and gaa is None and gmu is None
        return
    np.testing.assert_array_equal(ga, wa)
    v = rng.normal(size=prob.amp_alpha.shape)
    fd = (f(amp_alpha=prob.amp_alpha + h * v) - f(amp_alpha=prob.amp_alpha - h * v)) / (2 * h)
    assert fd == pytest.approx(np.vdot(gaa, v), rel=1e-6)
    fd = (f(mu_alpha=prob.mu_alpha + h) - f(mu_alpha=prob.mu_alpha - h)) / (2 * h)
    assert fd == pytest.approx(gmu, rel=1e-6)


def test_metric_is_gauss_newton_of_the_response():
    """v^T M v = sum (J v)^2 / lam with J v from central differences."""
    prob, xi = make_problem(H=32, n=4, margin=4, E=2, M=1, seed=4)
    rng = np.random.default_rng(1)
    v = (rng.normal(size=xi[0].shape), rng.normal(size=xi[1].shape))
    w = (rng.normal(size=xi[0].shape), rng.normal(size=xi[1].shape))
    Mv = om.metric_product(prob, xi[0], xi[1], *v)
    Mw = om.metric_product(prob, xi[0], xi[1], *w)
    # symmetric
    a = np.vdot(Mv[0], w[0]) + np.vdot(Mv[1], w[1])
    b = np.vdot(Mw[0], v[0]) + np.vdot(Mw[1], v[1])
    assert a == pytest.approx(b, rel=1e-10)
    h = 1e-6
    lam_p = om.response(prob, om.sky(prob, xi[0] + h * v[0], xi[1] + h * v[1]))
    lam_m = om.response(prob, om.sky(prob, xi[0] - h * v[0], xi[1] - h * v[1]))
    lam0 = om.response(prob, om.sky(prob, *xi))
    quad = sum(np.sum(((p - q) / (2 * h))[mod.keep] ** 2 / l0[mod.keep])
               for p, q, l0, mod in zip(lam_p, lam_m, lam0, prob.modules))
    vMv = np.vdot(Mv[0], v[0]) + np.vdot(Mv[1], v[1])
    assert vMv == pytest.approx(quad, rel=1e-6)


@pytest.mark.parametrize("slope,fluct,E", [(-4.0, 1.0, 1), (-3.0, 0.7, 3)])
def test_powerlaw_hyperparameter_gradients(ja, slope, fluct, E):
    """powerlaw_amplitude_grad chained to the oracle's amplitude-field
    gradient against central differences of the NLL in the power-law
    slope and fluctuations (the amplitude field rebuilt each time)."""
    from dataclasses import replace
    N = 32
    prob, xi = make_problem(H=32, n=4, margin=4, E=E, M=1, seed=41)

    def with_hyper(s, f):
        a = ja.powerlaw_amplitude(N, slope=s, fluctuations=f)
        return replace(prob, amp_ref=a, amp_alpha=None if prob.amp_alpha is None else 0.5 * a)

    p0 = with_hyper(slope, fluct)
    _, _, _, gar, gaa, _, _ = om.nll_and_grad_hyper(p0, *xi)
    # amp_alpha = amp_ref / 2 shares the hyperparameters: its gradient adds at half weight
    g_amp = gar if gaa is None else gar + 0.5 * gaa
    g_s, g_f = ja.powerlaw_amplitude_grad(N, g_amp, slope=slope, fluctuations=fluct)
    h = 1e-5
    fd_s = (om.nll(with_hyper(slope + h, fluct), *xi) -
            om.nll(with_hyper(slope - h, fluct), *xi)) / (2 * h)
    fd_f = (om.nll(with_hyper(slope, fluct + h), *xi) -
            om.nll(with_hyper(slope, fluct - h), *xi)) / (2 * h)
    assert g_s == pytest.approx(fd_s, rel=1e-6)
    assert g_f == pytest.approx(fd_f, rel=1e-6)
    # closed form of the fluctuations gradient: d amp / d f = amp / f
    amp = ja.powerlaw_amplitude(N, slope=slope, fluctuations=fluct)
    assert g_f == pytest.approx(float(np.sum(g_amp * amp)) / fluct, rel=1e-12)
    with pytest.raises(ValueError):
        ja.powerlaw_amplitude_grad(N, np.zeros((N, N + 1)))
