"""The hand-written adjoint chain of the fused model against central finite
differences of its own NLL, in float64 on the GPU (step 1e-6): the
directional derivative g . v must match (NLL(xi + e v) - NLL(xi - e v)) / 2e.
The 1024^2 case is the plain-vs-adjoint geometry (one energy, 8 x 8 patches,
margin 64, L = 384) with its Gaussian+Moffat PSFs (beta 2.5, core sigma
1-4 px), exposure ramp 5e2-2e3 s and 5 % mask."""
import numpy as np
import pytest
import torch

from tests.problems import make_problem, to_product

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cfg", [
    dict(H=64, n=4, margin=8, E=2, M=1, seed=71),
    dict(H=1024, n=8, margin=64, E=1, M=1, seed=72, expo_range=(5e2, 2e3), mask_frac=0.05,
         moffat=2.5, sigma=(1.0, 4.0)),
])
def test_adjoint_matches_central_differences(ja, cfg):
    prob, xi = make_problem(**cfg)
    model = to_product(ja, prob, dtype=torch.float64)
    dev = lambda a: None if a is None else torch.as_tensor(a, device="cuda")
    xr, xa = dev(xi[0]), dev(xi[1])
    _, g_r, g_a = model.nll_and_grad(xr, xa)
    rng = np.random.default_rng(cfg["seed"])
    vr = dev(rng.normal(size=xi[0].shape))
    va = dev(rng.normal(size=xi[1].shape)) if xi[1] is not None else None
    gv = float((g_r.double() * vr).sum())
    if va is not None:
        gv += float((g_a.double() * va).sum())
    eps = 1e-6
    plus = model.nll_and_grad(xr + eps * vr, None if xa is None else xa + eps * va)[0]
    minus = model.nll_and_grad(xr - eps * vr, None if xa is None else xa - eps * va)[0]
    fd = (plus - minus) / (2 * eps)
    print(f"H={cfg['H']}: g.v {gv:.12e} central difference {fd:.12e} rel {abs(fd - gv) / abs(gv):.2e}")
    assert abs(fd - gv) <= 1e-6 * abs(gv)
