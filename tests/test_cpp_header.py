"""CPU: the C++ corelib mirror header compiles on its own and the test
program that drives it builds and links against libhasty_hip.so."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_header_compiles_standalone(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text('#include <hasty/corelib.hpp>\nint main() { return 0; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-D__HIP_PLATFORM_AMD__",
                        "-I", os.path.join(ROOT, "include"), "-I", "/opt/rocm/include", str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_cpp_test_program_links():
    from hastycucompute_amd import build
    from hastycucompute_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        build.build()
    path = build.build_cpp_test()
    assert os.access(path, os.X_OK)
