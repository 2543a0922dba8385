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


def _uneven_models(ja, dtype):
    """Three observations on uneven grids: 64^2 x 2E (spectral), 128^2 with
    two modules (one energy, batched 2 samples), 96^2 x 2E on a 128^2 GP."""
    cfgs = [(dict(H=64, n=4, margin=8, E=2, M=1, seed=31), 1),
            (dict(H=128, n=4, margin=16, E=1, M=2, seed=32), 2),
            (dict(H=96, n=4, margin=24, E=2, M=1, seed=33, N=128), 1)]
    out = []
    for cfg, ms in cfgs:
        prob, xi = make_problem(**cfg)
        out.append((prob, xi, to_product(ja, prob, dtype, max_samples=ms)))
    return out


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_batch_graph_of_uneven_grids_matches_eager(ja, dtype):
    """jubik_graph_capture_batch: the evaluations of three models on uneven
    grids as parallel branches of one graph; every replay equals eager
    evaluations bit for bit, and launches set the single evaluations'
    metric base points."""
    models = _uneven_models(ja, dtype)
    samples = [0, 2, 0]
    graph = ja.capture_grad_batch([(m, S) for (_, _, m), S in zip(models, samples)])
    rng = np.random.default_rng(11)
    for _ in range(3):
        pos = []
        for (prob, xi, m), S, slot in zip(models, samples, graph.slots):
            lead = (S,) if S else ()
            pr = xi[0] + 0.05 * rng.normal(size=lead + xi[0].shape)
            pa = xi[1] + 0.05 * rng.normal(size=lead + xi[1].shape) if xi[1] is not None else None
            slot.xi_ref.copy_(torch.as_tensor(pr))
            if pa is not None:
                slot.xi_alpha.copy_(torch.as_tensor(pa))
            pos.append((pr, pa))
        graph.launch()
        torch.cuda.synchronize()
        for (prob, xi, m), S, slot, (pr, pa) in zip(models, samples, graph.slots, pos):
            if S:
                nll_e, gr_e, ga_e = m.nll_and_grad_batch(pr, pa)
                assert np.array_equal(slot.nll.cpu().numpy(), nll_e)
            else:
                nll_e, gr_e, ga_e = m.nll_and_grad(pr, pa)
                assert float(slot.nll[0].item()) == nll_e
            assert np.array_equal(slot.g_ref.cpu().numpy(), gr_e)
            if ga_e is not None:
                assert np.array_equal(slot.g_alpha.cpu().numpy(), ga_e)
            else:
                assert slot.g_alpha is None
    # base points: fresh models, one launch, then the metric of each single
    # evaluation equals that of an eager evaluation at the same position
    fresh = _uneven_models(ja, dtype)
    g2 = ja.capture_grad_batch([(fresh[0][2], 0), (fresh[2][2], 0)])
    for (prob, xi, m), slot in zip((fresh[0], fresh[2]), g2.slots):
        slot.xi_ref.copy_(torch.as_tensor(xi[0]))
        slot.xi_alpha.copy_(torch.as_tensor(xi[1]))
    g2.launch()
    torch.cuda.synchronize()
    for (prob, xi, m), (_, _, m_eager) in zip((fresh[0], fresh[2]), (models[0], models[2])):
        v = np.random.default_rng(3).normal(size=(2,) + xi[0].shape)
        got = m.metric(v[0], v[1])
        m_eager.nll_and_grad(xi[0], xi[1])
        want = m_eager.metric(v[0], v[1])
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])


def test_batch_graph_rejects_shared_workspaces(ja):
    models = _uneven_models(ja, torch.float32)
    m0 = models[0][2]
    with pytest.raises(ValueError):
        ja.capture_grad_batch([(m0, 0), (m0, 0)])
    twin = ja.XrayModel(m0.sky, m0.psf, 1.0, m0.exposure, m0.counts)
    with pytest.raises(ValueError):
        ja.capture_grad_batch([(m0, 0), (twin, 0)])
    with pytest.raises(ValueError):
        ja.capture_grad_batch([(models[1][2], 3)])  # above max_samples


def test_two_models_on_two_streams_match_serial_results(ja):
    """Entry points run on the caller's stream: two models evaluated on two
    side streams at once give the results of serial evaluation, bit for bit."""
    models = _uneven_models(ja, torch.float32)
    (pa, xa, ma), (pb, xb, mb) = models[0], models[2]
    want_a = ma.nll_and_grad(xa[0], xa[1])
    want_b = mb.nll_and_grad(xb[0], xb[1])
    torch.cuda.synchronize()
    dev = lambda a: torch.as_tensor(a, dtype=torch.float32, device="cuda")
    ins_a, ins_b = (dev(xa[0]), dev(xa[1])), (dev(xb[0]), dev(xb[1]))
    outs_a, outs_b = ma._grads(), mb._grads()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    s1.wait_stream(torch.cuda.current_stream())
    s2.wait_stream(torch.cuda.current_stream())
    for _ in range(3):
        with torch.cuda.stream(s1):
            na, ga_r, ga_a = ma.nll_and_grad_device(*ins_a, out=outs_a)
        with torch.cuda.stream(s2):
            nb, gb_r, gb_a = mb.nll_and_grad_device(*ins_b, out=outs_b)
        torch.cuda.synchronize()
        assert float(na.item()) == want_a[0] and float(nb.item()) == want_b[0]
        assert np.array_equal(ga_r.cpu().numpy(), want_a[1])
        assert np.array_equal(ga_a.cpu().numpy(), want_a[2])
        assert np.array_equal(gb_r.cpu().numpy(), want_b[1])
        assert np.array_equal(gb_a.cpu().numpy(), want_b[2])
