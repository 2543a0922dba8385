"""Column-tile autotuning of the linpatch column passes: every tile gives
the same results to rounding (the same butterflies run per column, only
their grouping into work-groups and the compiler's contraction of
multiply-adds change: a few ulp of the largest output), the tuner returns
a compiled tile with its timings, and a tile cache round-trips."""
import json

import numpy as np
import pytest
import torch

from tests.problems import make_problem, psf_windows, to_product

pytestmark = pytest.mark.gpu


def _operator(ja, H, n, m, E, dtype, seed):
    rng = np.random.default_rng(seed)
    op = ja.LinpatchOperator((E, H, H), n, m, dtype=dtype)
    op.set_kernel(psf_windows(rng, n, E, 2 * (H // n)), layout="window")
    x = torch.as_tensor(rng.normal(size=(E, H, H)), dtype=dtype, device="cuda")
    gy = torch.as_tensor(rng.normal(size=(E, H, H)), dtype=dtype, device="cuda")
    return op, x, gy


# Largest difference allowed, relative to the largest output: 4 ulp.
ULPS = {torch.float32: 4 * 2.0 ** -23, torch.float64: 4 * 2.0 ** -52}


def _close(a, b, dtype):
    return float((a - b).abs().max()) <= ULPS[dtype] * float(b.abs().max())


def _compiled(op, col, row):
    try:
        op.set_tile(col, row)
        return True
    except ValueError:  # a row tile under one row pair is not compiled in
        return False


# One geometry per compiled FFT length L = 2 (H / n) + 2 m.
@pytest.mark.parametrize("H,n,m,E", [(8, 2, 4, 1),      # L = 16
                                     (16, 2, 4, 1),     # 24
                                     (16, 2, 8, 2),     # 32
                                     (32, 2, 8, 1),     # 48
                                     (32, 2, 16, 2),    # 64
                                     (128, 4, 16, 1),   # 96
                                     (64, 2, 32, 1),    # 128
                                     (128, 2, 32, 1),   # 192
                                     (128, 2, 48, 1),   # 224
                                     (128, 2, 64, 2),   # 256
                                     (303, 3, 61, 1),   # 324
                                     (256, 2, 64, 2)])  # 384
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_every_tile_agrees_to_rounding(ja, H, n, m, E, dtype):
    op, x, gy = _operator(ja, H, n, m, E, dtype, seed=H + n)
    ref_y, ref_g = op.forward(x).clone(), op.adjoint(gy).clone()
    ran = 0
    for col, row in ((2, 1), (4, 1), (1, 2), (1, 4), (4, 4)):
        if not _compiled(op, col, row):
            continue
        ran += 1
        assert _close(op.forward(x), ref_y, dtype), f"forward, tiles {col}, {row}"
        assert _close(op.adjoint(gy), ref_g, dtype), f"adjoint, tiles {col}, {row}"
    assert ran >= 3
    # back to the default tiles: bit for bit the first result again
    op.set_tile(1, 1)
    assert torch.equal(op.forward(x), ref_y) and torch.equal(op.adjoint(gy), ref_g)


def test_autotune_picks_a_compiled_tile(ja, tmp_path):
    op, x, gy = _operator(ja, 128, 4, 16, 1, torch.float32, seed=5)
    ref_y = op.forward(x).clone()
    res = op.autotune(reps=5)
    assert res["col_div"] in (1, 2, 4) and res["row_div"] in (1, 2, 4) and not res["cached"]
    cols, rows = res["ms"]["col"], res["ms"]["row"]
    assert set(cols) == {1, 2, 4} and 1 in rows and all(t > 0 for t in (*cols.values(),
                                                                           *rows.values()))
    assert cols[res["col_div"]] == min(cols.values())
    assert rows[res["row_div"]] == min(rows.values())
    assert (op.col_div, op.row_div) == (res["col_div"], res["row_div"])
    assert _close(op.forward(x), ref_y, torch.float32)
    # the choice is stored and read back by a plan of the same geometry
    cache = tmp_path / "tiles.json"
    first = op.autotune(reps=5, cache=cache)
    table = json.loads(cache.read_text())
    assert table[op.tune_key()] == [first["col_div"], first["row_div"]]
    table[op.tune_key()] = [4, 2]
    cache.write_text(json.dumps(table))
    op2, _, _ = _operator(ja, 128, 4, 16, 1, torch.float32, seed=5)
    again = op2.autotune(cache=cache)
    assert again == {"col_div": 4, "row_div": 2, "ms": None, "cached": True}
    assert (op2.col_div, op2.row_div) == (4, 2)


def test_invalid_tile_is_rejected(ja):
    op, _, _ = _operator(ja, 128, 4, 16, 1, torch.float32, seed=6)
    for d in (0, 3, 8, -1):
        with pytest.raises(ValueError):
            op.set_tile(d)
        with pytest.raises(ValueError):
            op.set_tile(1, d)


@pytest.mark.parametrize("E,M", [(2, 1), (1, 3)])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_model_tiles_agree(ja, dtype, E, M):
    prob, _ = make_problem(H=256, n=4, margin=32, E=E, M=M, seed=81)
    model = to_product(ja, prob, dtype=dtype, max_samples=4)
    rng = np.random.default_rng(8)
    N = model.sky.N
    xr = torch.as_tensor(rng.normal(size=(4, N, N)), dtype=dtype, device="cuda")
    xa = torch.as_tensor(rng.normal(size=(4, N, N)), dtype=dtype, device="cuda")
    want = model.nll_and_grad_batch(xr, xa)
    res = model.autotune(samples=4, reps=3)
    assert res["col_div"] in (1, 2, 4) and set(res["ms"]["col"]) == {1, 2, 4}
    # gradients to the bar of the FP32 kernels against FP64 (1e-5), tighter
    # in FP64; the default tiles again give bit-identical results
    tol = 1e-6 if dtype == torch.float32 else 1e-13
    for col, row in ((2, 0), (4, 0), (1, 2), (1, 4), (1, 1)):
        try:
            model.set_tile(col, row)
        except ValueError:
            continue
        got = model.nll_and_grad_batch(xr, xa)
        assert np.allclose(got[0], want[0], rtol=tol, atol=0)
        for g, w in zip(got[1:], want[1:]):
            assert (g is None) == (w is None)
            if w is not None:
                assert float((g - w).norm() / w.norm()) <= tol
    assert np.array_equal(got[0], want[0]) and torch.equal(got[1], want[1])
    assert (got[2] is None and want[2] is None) or torch.equal(got[2], want[2])
