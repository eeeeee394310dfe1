"""CPU: the host thread pool with per-thread storage (include/hasty/threading.hpp,
the reference's threading.cppm:15-395), driven by tests/cpp/test_threading.cpp,
built with g++ and again with ThreadSanitizer (host code only).  The sanitizer
build uses ROCm's clang: GCC 11's ThreadSanitizer runtime does not intercept
pthread_cond_clockwait, which libstdc++ uses for steady-clock waits, and then
reports false double locks and races."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "test_threading.cpp")


CLANG = "/opt/rocm/llvm/bin/clang++"


def _build_and_run(tmp_path, extra, cxx="g++"):
    exe = tmp_path / "threading"
    r = subprocess.run([cxx, "-std=c++17", "-O1", "-g", "-Wall", "-Wextra", "-Werror", "-pthread", *extra,
                        "-I", os.path.join(ROOT, "include"), SRC, "-o", str(exe)],
                       capture_output=True, text=True)
    if r.returncode != 0:
        return r
    return subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)


def test_threading_pool(tmp_path):
    r = _build_and_run(tmp_path, [])
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL OK" in r.stdout


def test_threading_pool_under_thread_sanitizer(tmp_path):
    if not os.path.exists(CLANG):
        pytest.skip("no clang++ with a ThreadSanitizer runtime")
    probe = subprocess.run([CLANG, "-fsanitize=thread", "-x", "c++", "-", "-o", str(tmp_path / "p")],
                           input="int main(){return 0;}", capture_output=True, text=True)
    if probe.returncode != 0:
        pytest.skip("clang++ has no ThreadSanitizer runtime here")
    r = _build_and_run(tmp_path, ["-fsanitize=thread"], CLANG)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL OK" in r.stdout
    assert "WARNING: ThreadSanitizer" not in r.stderr, r.stderr
