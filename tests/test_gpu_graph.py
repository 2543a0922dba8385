"""The fused gradient evaluation replayed from a captured HIP graph.

The library allocates nothing, synchronises nothing and records no memset
inside an evaluation, so a caller can capture it.  Every replay (with new
latent inputs copied into the captured buffers, and eager work and
synchronisations in between) must equal an eager evaluation bit for bit.
"""
import numpy as np
import pytest
import torch

from tests.problems import make_problem, to_product

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_gradient_graph_replays_match_eager(ja, dtype):
    prob, xi = make_problem(H=64, n=4, margin=8, E=2, M=2, seed=21)
    model = to_product(ja, prob, dtype)
    xr = torch.as_tensor(xi[0], dtype=dtype, device="cuda")
    xa = torch.as_tensor(xi[1], dtype=dtype, device="cuda")
    out = model._grads()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        model.nll_and_grad_device(xr, xa, out=out)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        nll_t, gr_t, ga_t = model.nll_and_grad_device(xr, xa, out=out)
    rng = np.random.default_rng(5)
    for _ in range(3):
        pr = xi[0] + 0.05 * rng.normal(size=xi[0].shape)
        pa = xi[1] + 0.05 * rng.normal(size=xi[1].shape)
        xr.copy_(torch.as_tensor(pr))
        xa.copy_(torch.as_tensor(pa))
        graph.replay()
        torch.cuda.synchronize()
        got = (float(nll_t.item()), gr_t.cpu().numpy().copy(), ga_t.cpu().numpy().copy())
        nll_e, gr_e, ga_e = model.nll_and_grad(torch.as_tensor(pr, dtype=dtype, device="cuda"),
                                               torch.as_tensor(pa, dtype=dtype, device="cuda"))
        torch.cuda.synchronize()
        assert got[0] == nll_e
        assert np.array_equal(got[1], gr_e.cpu().numpy())
        assert np.array_equal(got[2], ga_e.cpu().numpy())
