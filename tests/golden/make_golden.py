"""Write the self-generated regression vectors of the oracle.

These are outputs of this repository's float64 oracle (not of J-UBIK, which
cannot run here): small seeded problems with their NLL, gradient and a
metric product.  tests/test_golden.py checks that the oracle still
reproduces them and the GPU tests compare the kernels against them.
Run from the repository root:  python tests/golden/make_golden.py
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import model as om  # noqa: E402
from tests.problems import make_problem  # noqa: E402

CASES = {
    "single_energy_32": dict(H=32, n=4, margin=4, E=1, M=1, seed=21),
    "spectral_64": dict(H=64, n=4, margin=8, E=2, M=1, seed=22),
}


def main():
    for name, cfg in CASES.items():
        prob, xi = make_problem(**cfg)
        nll, (gr, ga) = om.nll_and_grad(prob, *xi)
        rng = np.random.default_rng(99)
        v = (rng.normal(size=xi[0].shape), None if xi[1] is None else rng.normal(size=xi[1].shape))
        mr, ma = om.metric_product(prob, xi[0], xi[1], *v)
        mod = prob.modules[0]
        out = dict(cfg=np.array(repr(cfg)), xi_ref=xi[0], amp_ref=prob.amp_ref, window=mod.window,
                   exposure=mod.exposure, keep=mod.keep, data=mod.data, dnu=prob.dnu,
                   zm=prob.zm, mu=prob.mu_alpha, nll=nll, g_ref=gr, v_ref=v[0], m_ref=mr)
        if xi[1] is not None:
            out.update(xi_alpha=xi[1], amp_alpha=prob.amp_alpha, g_alpha=ga, v_alpha=v[1],
                       m_alpha=ma)
        np.savez_compressed(Path(__file__).with_name(f"{name}.npz"), **out)
        print(name, nll)


if __name__ == "__main__":
    main()
