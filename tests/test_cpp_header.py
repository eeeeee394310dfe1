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


def test_binary_tensor_io_host_only(tmp_path):
    """export/import_binary_tensor of the C++ mirror: host-only round trip,
    raw bytes (no header) readable by the Python mirror, and its errors."""
    src = tmp_path / "io.cpp"
    out = tmp_path / "blk.htc"
    src.write_text(r'''
#include <hasty/corelib.hpp>
#include <cstdio>
using namespace hasty;
int main(int argc, char** argv) {
    tensor<cpu_t, c64_t, 2> t({3, 4});
    for (int i = 0; i < 12; ++i) t.vec()[size_t(i)] = c64_t(float(i), -0.5f * float(i));
    export_binary_tensor(t, argv[1]);
    auto u = import_binary_tensor<c64_t, 2>(argv[1], {3, 4});
    for (int i = 0; i < 12; ++i) if (u.vec()[size_t(i)] != t.vec()[size_t(i)]) return 2;
    try { (void)import_binary_tensor<c64_t, 2>(argv[1], {4, 4}); return 3; }
    catch (const std::runtime_error& e) { if (std::string(e.what()) != "Exception opening/reading file") return 4; }
    try { (void)import_binary_tensor<c64_t, 1>(std::string(argv[1]) + ".none", {1}); return 5; }
    catch (const std::runtime_error& e) { if (std::string(e.what()) != "Could not open file for reading") return 6; }
    return 0;
}
''')
    exe = tmp_path / "io"
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"),
                        "-I", "/opt/rocm/include", str(src), "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe), str(out)], capture_output=True, text=True)
    assert r.returncode == 0, (r.returncode, r.stderr)
    import torch
    from hastycucompute_amd.io import import_binary_tensor
    t = import_binary_tensor(out, (3, 4), torch.complex64)
    ref = torch.complex(torch.arange(12.0), -0.5 * torch.arange(12.0)).reshape(3, 4)
    assert torch.equal(t, ref)


def test_cache_tensor_host_only(tmp_path):
    """cache_tensor of the C++ mirror (mod_tensor_caching.cppm): shape, shared
    copies, spill to cache_dir / "<hash>.htc" (the raw file the Python mirror
    reads), reload after free_cpu, and the reference's error messages."""
    src = tmp_path / "ct.cpp"
    src.write_text(r'''
#include <hasty/corelib.hpp>
using namespace hasty;
int main(int argc, char** argv) {
    cache_dir = argv[1];
    tensor<cpu_t, c64_t, 2> t({2, 3});
    for (int i = 0; i < 6; ++i) t.vec()[size_t(i)] = c64_t(float(i), 1.5f * float(i));
    cache_tensor<c64_t, 2> c(std::move(t), 4242);
    if (c.shape<0>() != 2 || c.shape()[1] != 3) return 2;
    auto d = c.copy();
    c.cache();
    c.free_cpu();
    auto h = d.get<cpu_t>();                       // shared block: reloaded from the file
    for (int i = 0; i < 6; ++i) if (h.vec()[size_t(i)] != c64_t(float(i), 1.5f * float(i))) return 3;
    if (c.get_ptr<cpu_t>().get() != d.get_ptr<cpu_t>().get()) return 4;
    c.uncache();
    if (c.get<cpu_t>().vec()[5] != c64_t(5.f, 7.5f)) return 5;
    cache_tensor<c64_t, 2> empty;
    try { (void)empty.shape(); return 6; }
    catch (const std::runtime_error& e) { if (std::string(e.what()) != "shape: cache_tensor is not initialized") return 7; }
    cache_tensor<c64_t, 1> none(tensor<cpu_t, c64_t, 1>({4}), 77);
    try { none.free_cpu(); (void)none.get<cpu_t>(); return 8; }
    catch (const std::runtime_error& e) { if (std::string(e.what()) != "Could not open file for reading") return 9; }
    return 0;
}
''')
    exe = tmp_path / "ct"
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-D__HIP_PLATFORM_AMD__",
                        "-I", os.path.join(ROOT, "include"), "-I", "/opt/rocm/include", str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True)
    assert r.returncode == 0, (r.returncode, r.stderr)
    import torch
    from hastycucompute_amd.io import import_binary_tensor
    back = import_binary_tensor(tmp_path / "4242.htc", (2, 3), torch.complex64)
    assert torch.equal(back, torch.complex(torch.arange(6.0), 1.5 * torch.arange(6.0)).reshape(2, 3))
