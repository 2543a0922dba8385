"""Host-side logic without a GPU: the C ABI library loads and exports every
symbol include/jubik_hip.h declares, arguments are validated before any
device work (with the reference's exception types and messages), and the
compute path refuses to run without a HIP device instead of falling back."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "jubik_hip.h"


def _declared():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"\b(jubik_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = _declared()
    for must in ("jubik_linpatch_create", "jubik_linpatch_forward", "jubik_linpatch_adjoint",
                 "jubik_poisson_nll_grad", "jubik_gp_forward", "jubik_gp_vjp",
                 "jubik_model_grad", "jubik_model_metric", "jubik_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol(ja):
    from jubik_amd import _lib
    lib = _lib.load()
    for name in _declared():
        assert hasattr(lib, name), name
    assert set(_lib.SIGNATURES) | {"jubik_last_error", "jubik_convolve_work_bytes"} >= set(
        _declared())
    assert lib.jubik_abi_version() == 1


def test_invalid_arguments_fail_before_device_work(ja):
    from jubik_amd import _lib
    lib = _lib.load()
    plan = ctypes.c_void_p()
    # margin 0
    st = lib.jubik_linpatch_create(ctypes.byref(plan), 0, 64, 64, 1, 1, 4, 0, 1, 1)
    assert st == _lib.EINVAL and b"margin" in lib.jubik_last_error()
    # bad dtype
    st = lib.jubik_linpatch_create(ctypes.byref(plan), 7, 64, 64, 1, 1, 4, 8, 1, 1)
    assert st == _lib.EINVAL
    # no compiled FFT length: window 2 * 1024 needs at least 4096 (fold mode)
    st = lib.jubik_linpatch_create(ctypes.byref(plan), 0, 2048, 2048, 1, 1, 2, 9, 1, 1)
    assert st == _lib.EUNSUPPORTED and b"window 2048" in lib.jubik_last_error()
    # unequal strides
    st = lib.jubik_linpatch_create(ctypes.byref(plan), 0, 64, 96, 1, 1, 4, 8, 1, 1)
    assert st == _lib.EINVAL
    with pytest.raises(ValueError):
        _lib.check(_lib.EINVAL)
    with pytest.raises(NotImplementedError):
        _lib.check(_lib.EUNSUPPORTED)
    # NULL plan pointers
    assert lib.jubik_linpatch_forward(None, None, None, 1.0, None) == _lib.EINVAL
    assert lib.jubik_model_grad(None, None, None, None, None, 0.0, 0.0, None, None, None, None,
                                None) == _lib.EINVAL


def test_tuning_and_batch_graph_entry_points_validate_arguments(ja):
    from jubik_amd import _lib
    lib = _lib.load()
    col, row = ctypes.c_int(), ctypes.c_int()
    ms = (ctypes.c_double * 6)()
    assert lib.jubik_linpatch_set_tile(None, 1, 1) == _lib.EINVAL
    assert lib.jubik_linpatch_autotune(None, 5, ctypes.byref(col), ctypes.byref(row), ms,
                                       None) == _lib.EINVAL
    assert lib.jubik_model_set_tile(None, 0, 0) == _lib.EINVAL
    assert lib.jubik_model_autotune(None, 1, 5, ctypes.byref(col), ctypes.byref(row), ms,
                                    None) == _lib.EINVAL
    graph = ctypes.c_void_p()
    assert lib.jubik_graph_capture_batch(ctypes.byref(graph), None, 1) == _lib.EINVAL
    descs = (_lib.GradDesc * 1)()
    st = lib.jubik_graph_capture_batch(ctypes.byref(graph), ctypes.cast(descs, ctypes.c_void_p), 0)
    assert st == _lib.EINVAL and b"count" in lib.jubik_last_error()
    st = lib.jubik_graph_capture_batch(ctypes.byref(graph), ctypes.cast(descs, ctypes.c_void_p), 1)
    assert st == _lib.EINVAL and b"descriptor 0" in lib.jubik_last_error()


def test_newer_entry_points_validate_arguments(ja):
    """The transform switch, the general convolve, the point-source sky, the
    inverse-gamma prior and the geoVI entry points reject NULL handles and
    bad arguments with a status and a message, before any device work."""
    from jubik_amd import _lib
    lib = _lib.load()
    assert lib.jubik_linpatch_set_transform(None, 0) == _lib.EINVAL
    buf = (ctypes.c_float * 16)()
    # convolve: NULL operands, empty axes, bad dtype
    assert lib.jubik_convolve(0, None, buf, 1, 1, 4, 4, 1.0, buf, buf, None) == _lib.EINVAL
    st = lib.jubik_convolve(0, buf, buf, 1, 1, 0, 4, 1.0, buf, buf, None)
    assert st == _lib.EINVAL and b"empty" in lib.jubik_last_error()
    assert lib.jubik_convolve(5, buf, buf, 1, 1, 4, 4, 1.0, buf, buf, None) == _lib.EINVAL
    # point sources and the inverse-gamma prior
    assert lib.jubik_model_set_point_sources(None, 1, 1.0, 1.0, 0.0, 1.0) == _lib.EINVAL
    assert lib.jubik_model_grad_points(None, 1, *([None] * 4), 0.0, 0.0,
                                       *([None] * (len(_lib.SIGNATURES["jubik_model_grad_points"]) - 8))) == _lib.EINVAL
    st = lib.jubik_invgamma_prior(0, None, 0, -1.0, 1.0, None, None, None)
    assert st == _lib.EINVAL and b"a > 0" in lib.jubik_last_error()
    assert lib.jubik_invgamma_prior(0, None, 4, 1.0, 1.0, None, None, None) == _lib.EINVAL
    assert lib.jubik_invgamma_prior(3, buf, 4, 1.0, 1.0, buf, buf, None) == _lib.EINVAL
    # geoVI
    nll = ctypes.c_double()
    assert lib.jubik_model_geovi_point(None, None, None, None, None, 0.0, 0.0, None, None,
                                       ctypes.byref(nll)) == _lib.EINVAL
    assert lib.jubik_model_geovi_apply(None, 0, *([None] * 8)) == _lib.EINVAL
    nargs = len(_lib.SIGNATURES["jubik_model_geovi_solve"])
    assert lib.jubik_model_geovi_solve(*([None] * 8), 10, 1e-6, 0.0, 1,
                                       *([None] * (nargs - 12))) == _lib.EINVAL


def test_grad_desc_layout_matches_the_header(ja, tmp_path):
    """The ctypes mirror of jubik_grad_desc has the C struct's size and
    field offsets (checked with the host C compiler)."""
    import shutil
    import subprocess
    from jubik_amd import _lib
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no host C compiler")
    fields = [f for f, _ in _lib.GradDesc._fields_]
    src = tmp_path / "layout.c"
    src.write_text("#include <stdio.h>\n#include <stddef.h>\n#include \"jubik_hip.h\"\n"
                   "int main(void) {\n  printf(\"%zu\\n\", sizeof(jubik_grad_desc));\n"
                   + "".join(f'  printf("%zu\\n", offsetof(jubik_grad_desc, {f}));\n'
                             for f in fields)
                   + "  return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c99", "-I", str(HEADER.parent), str(src), "-o", str(exe)],
                   check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True,
                                          text=True).stdout.split()]
    assert got[0] == ctypes.sizeof(_lib.GradDesc)
    assert got[1:] == [getattr(_lib.GradDesc, f).offset for f in fields]


def test_last_error_is_per_thread(ja):
    """jubik_last_error reports the calling thread's failure even when
    another thread failed in between."""
    import threading
    from jubik_amd import _lib
    lib = _lib.load()
    barrier = threading.Barrier(2)
    seen = {}

    def worker(name, margin):
        plan = ctypes.c_void_p()
        if name == "b":
            barrier.wait()  # fail after thread a
        H, n = (64, 4) if name == "a" else (2048, 2)
        st = lib.jubik_linpatch_create(ctypes.byref(plan), 0, H, H, 1, 1, n, margin, 1, 1)
        if name == "a":
            barrier.wait()
        barrier.wait()  # both have failed
        seen[name] = (st, lib.jubik_last_error())

    ts = [threading.Thread(target=worker, args=("a", 0)),
          threading.Thread(target=worker, args=("b", 9))]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=30)
    assert seen["a"][0] == _lib.EINVAL and b"margin" in seen["a"][1]
    assert seen["b"][0] == _lib.EUNSUPPORTED and b"window 2048" in seen["b"][1]


def test_tile_cache_round_trip(ja, tmp_path):
    import json
    from jubik_amd.convolve import tile_cache_get, tile_cache_put
    cache = tmp_path / "tiles.json"
    assert tile_cache_get(None, "k") is None
    assert tile_cache_get(cache, "k") is None
    tile_cache_put(None, "k", (2, 1))  # no cache: nothing written
    assert not cache.exists()
    tile_cache_put(cache, "gfx950/f32/L384", (2, 1))
    tile_cache_put(cache, "gfx950/f64/L96", (4, 2))
    tile_cache_put(cache, "gfx950/f32/L384", (1, 4))
    assert tile_cache_get(cache, "gfx950/f32/L384") == (1, 4)
    assert tile_cache_get(cache, "gfx950/f64/L96") == (4, 2)
    assert tile_cache_get(cache, "other") is None
    assert json.loads(cache.read_text()) == {"gfx950/f32/L384": [1, 4], "gfx950/f64/L96": [4, 2]}


def test_python_api_errors_match_reference(ja):
    x = np.ones((1, 16, 16))
    with pytest.raises(ValueError, match="domain has to be an instance of jubik.Domain"):
        ja.linpatch_convolve(x, ((1, 16, 16), (1, 1, 1)), np.ones((16, 1, 16, 16)), 4, 2)
    with pytest.raises(ValueError, match="exposure_cut should be non-negative or None!"):
        ja.build_exposure_function(np.ones((1, 4, 4)), -1)
    with pytest.raises(ValueError, match="threshold should be positive!"):
        ja.build_readout_function(np.ones((1, 1, 4, 4)), -1)
    with pytest.raises(ValueError, match="length of keys should match"):
        ja.build_readout_function(np.ones((2, 1, 4, 4)), None, ("a",))
    with pytest.raises(ValueError, match="flags should have shape"):
        ja.build_readout_function(np.ones((1, 4, 4)), None, ("a",))(np.ones((1, 1, 4, 4)))
    with pytest.raises(ValueError, match="input should have shape"):
        ja.build_readout_function(np.ones((1, 1, 4, 4)), None)(np.ones((1, 4, 4)))


def test_builders_mutate_callers_arrays_like_reference(ja):
    e = np.array([[[[1.0, 5.0], [30.0, 2.0]]]])
    ja.build_exposure_function(e, 3.0)
    np.testing.assert_array_equal(e, [[[[0.0, 5.0], [30.0, 0.0]]]])
    f = np.array([[[[1.0, 5.0], [30.0, 2.0]]]])
    ro = ja.build_readout_function(f, 3.0)
    np.testing.assert_array_equal(f, [[[[0.0, 5.0], [30.0, 0.0]]]])
    assert ro.keys == ("masked input",)


def test_domain_duck_typing(ja):
    from typing import NamedTuple

    from jubik_amd.data import is_domain

    class Domain(NamedTuple):
        shape: tuple
        distances: tuple

    assert is_domain(ja.Domain((1,), (1.0,)))
    assert is_domain(Domain((1,), (1.0,)))
    assert not is_domain(((1,), (1.0,)))


def test_dense_counts_wire_format(ja):
    keep = np.zeros((1, 1, 2, 3), dtype=bool)
    keep[0, 0, 0, 1] = keep[0, 0, 1, 2] = True
    out = ja.dense_counts([np.array([4, 7])], keep)
    np.testing.assert_array_equal(out, [[[[-1, 4, -1], [-1, -1, 7]]]])
    with pytest.raises(ValueError):
        ja.dense_counts([np.array([4])], keep)
    with pytest.raises(ValueError):
        ja.dense_counts([np.array([4, -1])], keep)


def test_margin_zero_is_empty_without_device(ja):
    x = np.ones((1, 16, 16))
    out = ja.linpatch_convolve(x, ja.Domain((1, 16, 16), (1, 1, 1)), np.ones((16, 1, 16, 16)),
                               4, 0)
    assert out.shape == (1, 0, 0)


def test_compute_path_requires_a_device(ja):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a HIP device is visible")
    x = np.ones((1, 16, 16))
    with pytest.raises(RuntimeError, match="no HIP\\s+device"):
        ja.linpatch_convolve(x, ja.Domain((1, 16, 16), (1, 1, 1)), np.ones((16, 1, 16, 16)), 4, 2)
    with pytest.raises(RuntimeError):
        ja.build_exposure_function(np.ones((1, 4, 4)))(np.ones((4, 4)))


def test_integrate_matches_reference_fixture(ja):
    assert ja.integrate(np.ones((3, 4)), ja.Domain((3, 4), (0.5, 0.25)), (0, 1)) == 1.5


def test_save_mock_data_writes_the_reference_pickles(ja, tmp_path):
    """data.py:146-149: the data Vector's dict and the position's dict are
    pickled under res_dir with the file names of file_info."""
    import pickle
    data = ja.Vector({"tm_1": np.array([0, 3, 7]), "tm_2": np.array([1, 1])})
    pos = {"xi_ref": np.arange(6.0).reshape(2, 3), "xi_alpha": np.ones((2, 3))}
    info = {"res_dir": str(tmp_path / "mock"), "data_dict": "data.pkl", "pos_dict": "pos.pkl"}
    dpath, ppath = ja.save_mock_data(data, pos, info)
    assert dpath == str(tmp_path / "mock" / "data.pkl") and ppath.endswith("pos.pkl")
    with open(dpath, "rb") as f:
        d = pickle.load(f)
    with open(ppath, "rb") as f:
        p = pickle.load(f)
    assert set(d) == {"tm_1", "tm_2"} and np.array_equal(d["tm_1"], [0, 3, 7])
    assert np.array_equal(p["xi_ref"], pos["xi_ref"]) and np.array_equal(p["xi_alpha"], np.ones((2, 3)))


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus N > 1 exits with status 2 and says
    so, before touching a device (also with --sweep)."""
    import subprocess
    import sys
    from pathlib import Path
    root = Path(__file__).resolve().parent.parent
    for extra in ([], ["--sweep"]):
        res = subprocess.run([sys.executable, str(root / "bench.py"), "--gpus", "2", *extra],
                             capture_output=True, text=True, timeout=120, cwd=root)
        assert res.returncode == 2
        assert "measures one GPU" in res.stderr
