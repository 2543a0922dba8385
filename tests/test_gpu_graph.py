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


@pytest.mark.parametrize("samples", [0, 3])
def test_library_graph_matches_eager(ja, samples):
    """jubik_model_grad_graph: a graph captured by the library itself,
    replayed with new positions, equals eager evaluations bit for bit."""
    prob, xi = make_problem(H=64, n=4, margin=8, E=2, M=2, seed=22)
    model = to_product(ja, prob, torch.float32, max_samples=max(samples, 1))
    graph = model.capture_grad(samples)
    rng = np.random.default_rng(9)
    lead = (samples,) if samples else ()
    for _ in range(3):
        pr = xi[0] + 0.05 * rng.normal(size=lead + xi[0].shape)
        pa = xi[1] + 0.05 * rng.normal(size=lead + xi[1].shape)
        graph.xi_ref.copy_(torch.as_tensor(pr))
        graph.xi_alpha.copy_(torch.as_tensor(pa))
        graph.launch()
        torch.cuda.synchronize()
        if samples:
            nll_e, gr_e, ga_e = model.nll_and_grad_batch(pr, pa)
            assert np.array_equal(graph.nll.cpu().numpy(), nll_e)
        else:
            nll_e, gr_e, ga_e = model.nll_and_grad(pr, pa)
            assert float(graph.nll[0].item()) == nll_e
        assert np.array_equal(graph.g_ref.cpu().numpy(), gr_e)
        assert np.array_equal(graph.g_alpha.cpu().numpy(), ga_e)
    if samples == 0:
        # a launch of the single graph sets the metric base point
        model2 = to_product(ja, prob, torch.float32)
        g2 = model2.capture_grad(0)
        g2.xi_ref.copy_(torch.as_tensor(xi[0]))
        g2.xi_alpha.copy_(torch.as_tensor(xi[1]))
        g2.launch()
        v = rng.normal(size=xi[0].shape)
        m_graph = model2.metric(v, v)
        model2.nll_and_grad(xi[0], xi[1])
        m_eager = model2.metric(v, v)
        assert np.array_equal(m_graph[0], m_eager[0])
