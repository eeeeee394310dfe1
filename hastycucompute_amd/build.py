"""Build the gfx950 shared library ``libhasty_hip.so`` in-tree with hipcc.

Usage: ``python -m hastycucompute_amd.build`` (or ``build()`` from Python).
Each translation unit is compiled in parallel to ``_build/*.o`` and linked
into ``hastycucompute_amd/libhasty_hip.so``.  No torch headers are involved:
the library's ABI is plain C.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libhasty_hip.so")
SOURCES = ["fft.hip", "fft_reg_f32.hip", "fft_reg_f64.hip", "nufft.hip", "nufft_kern_f32.hip",
           "nufft_kern_f64.hip", "toeplitz.hip", "toeplitz_fused.hip", "nufft_fused.hip", "cg.hip", "abi.cpp"]
ARCH = os.environ.get("HASTY_OFFLOAD_ARCH", "gfx950")
# Per-file flags.  The register FFT passes fold arithmetic on known-zero
# (padded) inputs away, which needs x + 0 -> x: signed zeros do not matter
# to an FFT, so those files are built without them.
EXTRA_FLAGS = {"fft_reg_f32.hip": ["-fno-signed-zeros"], "fft_reg_f64.hip": ["-fno-signed-zeros"],
               "toeplitz_fused.hip": ["-fno-signed-zeros"], "nufft_fused.hip": ["-fno-signed-zeros"]}


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def common_flags():
    return [
        "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
        # packed FP32 ops disabled for every device pass
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", os.path.join(ROOT, "include"),
        "-Wall", "-Wno-unused-result", "-Wno-unknown-pragmas",
    ]


def _compile(src: str) -> str:
    obj = os.path.join(OUT_DIR, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc()] + common_flags() + EXTRA_FLAGS.get(src, []) + ["-x", "hip", "-c", os.path.join(CSRC, src),
                                                                   "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile of {src} failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    warn = [ln for ln in r.stderr.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
    if warn and os.environ.get("HASTY_BUILD_VERBOSE"):
        print("\n".join(warn), file=sys.stderr)
    return obj


def build(jobs: int | None = None) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    jobs = jobs or min(len(SOURCES), max(1, (os.cpu_count() or 2)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_cpp_test()
    return LIB


CPP_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "test_corelib.cpp")
CPP_TEST_BIN = os.path.join(OUT_DIR, "test_corelib")


def build_cpp_test() -> str:
    """Compile the C++ corelib-mirror test program (host code only, g++)
    against the built library; it runs on a GPU from tests/test_gpu_cpp_api.py."""
    os.makedirs(OUT_DIR, exist_ok=True)
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-D__HIP_PLATFORM_AMD__",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(rocm, "include"),
           CPP_TEST_SRC, "-L", HERE, "-lhasty_hip", "-L", os.path.join(rocm, "lib"), "-lamdhip64",
           "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath," + os.path.join(rocm, "lib"),
           "-o", CPP_TEST_BIN + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"C++ test build failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(CPP_TEST_BIN + ".tmp", CPP_TEST_BIN)
    return CPP_TEST_BIN


if __name__ == "__main__":
    print(build())
