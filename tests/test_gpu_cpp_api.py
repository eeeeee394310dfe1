"""GPU: the C++ corelib mirror (include/hasty/corelib.hpp) driven by the
program tests/cpp/test_corelib.cpp, which checks nufft_plan type 1 / type 2,
nufft_backward_cpu_over_cuda, toeplitz_kernel, toeplitz_multiply,
fft::toeplitz_multiplication, the normal-operator family (also built from
cache_tensors), time-segmented SENSE, conjugate_gradient and max_eig
against a direct double-precision NDFT, and the reference's error messages."""
import os
import subprocess

import pytest


@pytest.mark.gpu
def test_cpp_corelib_mirror():
    from hastycucompute_amd import build
    if not os.path.exists(build.CPP_TEST_BIN):
        build.build_cpp_test()
    print(f"[cpp] running {build.CPP_TEST_BIN}", flush=True)
    r = subprocess.run([build.CPP_TEST_BIN], capture_output=True, text=True, timeout=110)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL OK" in r.stdout
