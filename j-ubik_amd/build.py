"""Build libjubik_hip.so for gfx950 with hipcc (in-tree, no JIT cache).

Each ``csrc/*.hip`` translation unit is compiled in parallel into
``_build/`` and linked into ``libjubik_hip.so`` next to this file.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
SRC = PKG / "csrc"
OUT = PKG / "_build"
LIB = PKG / "libjubik_hip.so"
ARCH = os.environ.get("JUBIK_OFFLOAD_ARCH", "gfx950")

# Packed FP32 instructions stay off: only scalar/float2 code is written and
# the compiler is told not to form v_pk_* FP32 operations from it.
CFLAGS = [
    "-std=c++17", "-O3", f"--offload-arch={ARCH}", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src):
    obj = OUT / (src.stem + ".o")
    cmd = [_hipcc(), *CFLAGS, "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{res.stderr[-6000:]}")
    return obj


def build(verbose=True):
    """Compile every HIP source and link the shared library; returns its path."""
    OUT.mkdir(exist_ok=True)
    sources = sorted(SRC.glob("*.hip"))
    jobs = min(len(sources), max(1, int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [_hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs),
           "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr[-6000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}", file=sys.stderr)
    return LIB


ASAN_LIB = OUT / "libjubik_hip_asan.so"


def _asan_runtime():
    """clang's AddressSanitizer runtime of the hipcc toolchain."""
    root = Path(_hipcc()).resolve().parent.parent
    for base in (root / "llvm" / "lib" / "clang", root / "lib" / "llvm" / "lib" / "clang"):
        for cand in sorted(base.glob("*/lib/linux/libclang_rt.asan-x86_64.so")):
            return cand
    raise RuntimeError("clang's AddressSanitizer runtime not found")


def build_asan(verbose=True):
    """Host-side AddressSanitizer variant of the library for the argument
    validation tests: capi.hip (the C ABI's checks and host logic) is
    compiled with ASan on the host side only, the device code and the other
    translation units are the regular objects of ``build``.  Returns
    (library path, ASan runtime to preload)."""
    objs = sorted(OUT.glob("*.o")) if OUT.exists() else []
    wanted = {src.stem for src in SRC.glob("*.hip")}
    objs = [o for o in objs if o.stem in wanted and o.stem != "capi"]
    if len(objs) != len(wanted) - 1:
        build(verbose=verbose)
        objs = [o for o in sorted(OUT.glob("*.o")) if o.stem in wanted and o.stem != "capi"]
    asan_obj = OUT / "capi_asan.o"
    cmd = [_hipcc(), *CFLAGS, "-g", "-Xarch_host", "-fsanitize=address", "-Xarch_host",
           "-fno-omit-frame-pointer", "-c", str(SRC / "capi.hip"), "-o", str(asan_obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc (host ASan) failed for capi.hip:\n{res.stderr[-6000:]}")
    tmp = ASAN_LIB.with_suffix(".so.tmp")
    cmd = [_hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs), str(asan_obj),
           "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link (host ASan) failed:\n{res.stderr[-6000:]}")
    os.replace(tmp, ASAN_LIB)
    if verbose:
        print(f"built {ASAN_LIB}", file=sys.stderr)
    return ASAN_LIB, _asan_runtime()


if __name__ == "__main__":
    build()
