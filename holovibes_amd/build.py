"""Builds holovibes_amd/libholo_hip.so for gfx950 with hipcc.

Every source under csrc/ is compiled as HIP (-x hip) into _build/, in
parallel, then linked into one shared library beside this file.  The twiddle
table is generated into _build/gen/ first.  Usage: python -m holovibes_amd.build
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
GEN = os.path.join(BUILD, "gen")
LIB = os.path.join(HERE, "libholo_hip.so")
# extra compiler flags (e.g. -D tile-shape overrides) for experiment builds
EXTRA = os.environ.get("HV_EXTRA_FLAGS", "").split()
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("HV_OFFLOAD_ARCH", "gfx950")

SOURCES = ["hv_capi.cpp", "hv_pipe.cpp", "hv_synth.cpp", "hv_spatial.hip", "hv_time.hip", "hv_bins.hip", "hv_render.hip",
           "hv_pca.hip", "hv_holo.cpp", "hv_input.hip",
           "hv_source.cpp", "hv_post.hip"]
HEADERS = ["hv_common.hpp", "hv_fft.hpp", "hv_kernels.hpp", os.path.join("..", "..", "include", "holo_hip.h")]

CFLAGS = [
    "-x", "hip", f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
    "-Wno-unused-variable", "-Wno-unused-but-set-variable",
    # no packed-FP32 instructions anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", GEN, "-I", CSRC,
]


def _gen_twiddles() -> None:
    os.makedirs(GEN, exist_ok=True)
    out = os.path.join(GEN, "hv_twiddles.inc")
    src = os.path.join(CSRC, "gen_twiddles.py")
    if not os.path.exists(out) or os.path.getmtime(out) < os.path.getmtime(src):
        subprocess.run([sys.executable, src, out], check=True)


def _compile(src: str, force: bool) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    deps = [os.path.join(CSRC, src)] + [os.path.join(CSRC, h) for h in HEADERS]
    if not force and os.path.exists(obj) and all(os.path.getmtime(obj) >= os.path.getmtime(d) for d in deps):
        return obj
    cmd = [HIPCC] + CFLAGS + EXTRA + ["-c", os.path.join(CSRC, src), "-o", obj + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    os.replace(obj + ".tmp", obj)
    return obj


def build(force: bool = True, jobs: int | None = None, out: str | None = None, build_dir: str | None = None) -> str:
    """Compiles every source (force: even when objects are newer) and links
    libholo_hip.so (or `out`).  Returns the library path."""
    global BUILD
    if build_dir:
        BUILD = build_dir
    os.makedirs(BUILD, exist_ok=True)
    _gen_twiddles()
    jobs = jobs or min(len(SOURCES), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), SOURCES))
    lib = out or LIB
    tmp = lib + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, lib)
    return lib


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--incremental", action="store_true")
    ap.add_argument("--out", default=None, help="library path (default holovibes_amd/libholo_hip.so)")
    ap.add_argument("--build-dir", default=None, help="object directory (default holovibes_amd/_build)")
    a = ap.parse_args()
    print(build(force=not a.incremental, out=a.out, build_dir=a.build_dir))
