"""The host-side argument validation of the C ABI under AddressSanitizer:
capi.hip is rebuilt with host ASan (libjubik_hip_asan.so, device code
unchanged) and the host tests (tests/test_host.py: every status code,
message, thread-local error text and tile-cache path, no device needed)
run against it in a subprocess with the ASan runtime preloaded.  Any heap
or stack error in the validation code ends that process with an ASan
report."""
import importlib.util
import os
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def _builder():
    spec = importlib.util.spec_from_file_location("jubik_amd_build",
                                                  ROOT / "j-ubik_amd" / "build.py")
    b = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(b)
    return b


def test_host_validation_under_asan():
    lib, runtime = _builder().build_asan(verbose=False)
    env = dict(os.environ)
    env.update({"JUBIK_HIP_LIB": str(lib), "LD_PRELOAD": str(runtime),
                "ASAN_OPTIONS": "detect_leaks=0:abort_on_error=0:exitcode=99"})
    res = subprocess.run([sys.executable, "-m", "pytest", str(ROOT / "tests" / "test_host.py"),
                          "-q", "-m", "not gpu", "-p", "no:cacheprovider"],
                         cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    out = res.stdout + res.stderr
    assert "AddressSanitizer" not in out, out[-4000:]
    assert res.returncode == 0, out[-4000:]
    assert " passed" in res.stdout
