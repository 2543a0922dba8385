"""Several telescope modules over a spectral sky: the fused NLL and gradient
of 7 modules x 3 energies (21 module-energy passes in one launch per pass),
each module with its own rotated PSF set and exposure map, FFT length
L = 324 (the 7-module workload: 10 x 10 patches of stride 102, margin 60).

At reduced size (3 x 3 patches of the same stride and margin) the GPU is
held to the float64 oracle; at the full 1024^2 size the FP32 gradient is
held to the FP64 build of the same kernels (the north-star gate of 1e-5)."""
import numpy as np
import pytest
import torch

from oracle import model as om
from tests.problems import make_problem, powerlaw_amp, psf_windows, to_product

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-10), (torch.float32, 1e-5)])
def test_seven_modules_spectral_vs_oracle(ja, dtype, tol):
    prob, xi = make_problem(H=306, n=3, margin=60, E=3, M=7, seed=3, N=512, sigma=(2.0, 6.0),
                            expo_range=(1e2, 3e3), mask_frac=0.08)
    want_nll, (wr, wa) = om.nll_and_grad(prob, *xi)
    model = to_product(ja, prob, dtype)
    assert model.psf.L == 324
    dev = lambda a: torch.as_tensor(a, dtype=dtype, device="cuda")
    nll, gr, ga = model.nll_and_grad(dev(xi[0]), dev(xi[1]))
    print(f"7 modules x 3E, L = 324: nll rel {abs(nll - want_nll) / abs(want_nll):.2e}, "
          f"grad rel {_rel(gr.double().cpu().numpy(), wr):.2e} / "
          f"{_rel(ga.double().cpu().numpy(), wa):.2e}")
    assert abs(nll - want_nll) <= tol * abs(want_nll)
    assert _rel(gr.double().cpu().numpy(), wr) <= tol
    assert _rel(ga.double().cpu().numpy(), wa) <= tol


def test_seven_modules_full_size_fp32_vs_fp64(ja):
    H, n, m, E, M = 1024, 10, 60, 3, 7
    rng = np.random.default_rng(3)
    amp = powerlaw_amp(H)
    dnu = np.log(np.geomspace(1.0, 4.0, E))
    w = 2 * (H // n)
    win = np.stack([psf_windows(rng, n, E, w, sigma=(2.0, 6.0), rotation=k * np.pi / 7)
                    for k in range(M)])
    ramp = np.linspace(1e2, 3e3, H)[None, None, None, :]
    expo = np.broadcast_to(ramp, (M, E, H, H)).copy()
    expo[rng.random(expo.shape) < 0.08] = 0.0
    keep = expo > 0
    xt, xat = rng.normal(size=(H, H)), rng.normal(size=(H, H))

    def build(dtype, counts):
        sky = ja.CorrelatedFieldSky(H, (H, H), amp, 0.5 * amp, zm=-2.0, mu=-1.0, dnu=dnu,
                                    dtype=dtype)
        psf = ja.LinpatchOperator((E, H, H), n, m, n_modules=M, dtype=dtype)
        psf.set_kernel(win, layout="window")
        return sky, psf, ja.XrayModel(sky, psf, 1.0, expo, counts)

    sky64, psf64, _ = build(torch.float64, np.zeros(keep.shape, np.int32))
    assert psf64.L == 324
    lam = psf64.forward(sky64(torch.as_tensor(xt, device="cuda"),
                              torch.as_tensor(xat, device="cuda"))) * torch.as_tensor(
        expo, device="cuda")
    counts = np.where(keep, ja.poisson_sample(lam, seed=3).cpu().numpy(), -1).astype(np.int32)
    del lam
    xi = (xt + 0.1 * rng.normal(size=(H, H)), xat + 0.1 * rng.normal(size=(H, H)))
    res = {}
    for dtype in (torch.float64, torch.float32):
        _, _, model = build(dtype, counts)
        nll, gr, ga = model.nll_and_grad(*[torch.as_tensor(v, dtype=dtype, device="cuda")
                                           for v in xi])
        res[dtype] = (nll, np.concatenate([gr.double().cpu().numpy().ravel(),
                                           ga.double().cpu().numpy().ravel()]))
        del model
    g32, g64 = res[torch.float32][1], res[torch.float64][1]
    err = _rel(g32, g64)
    nmax = float(np.max(np.abs(g32 - g64)) / np.max(np.abs(g64)))
    print(f"1024^2 x 3E x 7 modules: FP32 vs FP64 grad rel {err:.2e}, normalised max {nmax:.2e}")
    assert err <= 1e-5
    n64 = res[torch.float64][0]
    assert abs(res[torch.float32][0] - n64) <= 1e-6 * abs(n64)
