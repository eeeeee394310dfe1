"""HDF5 tensor files (host logic, no GPU): the reference's io/hdf5 interface
(``hdf5.cppm:14-19``) and its raw k-space loader (``test_interface.cpp:36-112``).

No HDF5 fixtures ship with the reference, so every file here is written by the
module itself and checked by exact round trips, by the reference's key layout
and scaling, and by its error behaviour.
"""
import os

import numpy as np
import pytest
import torch

from hastycucompute_amd import hdf5


@pytest.fixture
def tmpfile(tmp_path):
    return str(tmp_path / "t.h5")


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64, torch.complex64, torch.complex128])
@pytest.mark.parametrize("shape", [(7,), (3, 4, 5), (2, 1, 3, 2), (), (0, 3)])
def test_round_trip(tmpfile, dtype, shape):
    g = torch.Generator().manual_seed(1)
    t = torch.randn(shape, generator=g, dtype=dtype)
    hdf5.export_tensor(t, tmpfile, "/data/x")
    back = hdf5.import_tensor(tmpfile, "/data/x")
    assert back.dtype == dtype and tuple(back.shape) == shape
    assert torch.equal(back, t)


def test_export_overwrites_and_needs_contiguous(tmpfile):
    hdf5.export_tensor(torch.ones(3), tmpfile, "a")
    hdf5.export_tensor(torch.zeros(2), tmpfile, "b")
    assert list(hdf5.import_tensors(tmpfile)) == ["/b"]     # File::Overwrite
    with pytest.raises(RuntimeError, match="contiguous"):
        hdf5.export_tensor(torch.ones(4, 4).t(), tmpfile, "c")


@pytest.mark.parametrize("names", [("r", "i"), ("real", "imag")])
def test_complex_member_names(tmpfile, names):
    """Both member namings the reference reads (std::complex "r"/"i" and its
    Complex64 struct "real"/"imag") load as complex, real part first."""
    t = torch.tensor([1 + 2j, -3.5 + 0.25j], dtype=torch.complex64)
    hdf5.export_tensors({"/k": t}, tmpfile, complex_names=names)
    back = hdf5.import_tensor(tmpfile, "/k")
    assert torch.equal(back, t)


def test_import_tensors_walks_groups_and_matches_fully(tmpfile):
    ts = {"/Kdata/KX_E0": torch.arange(4.0), "/Kdata/KX_E10": torch.arange(2.0),
          "/Kdata/sub/KY_E0": torch.ones(3, dtype=torch.float64), "/top": torch.zeros(1)}
    hdf5.export_tensors(ts, tmpfile)
    everything = hdf5.import_tensors(tmpfile)
    assert sorted(everything) == sorted(ts)
    for k, v in ts.items():
        assert torch.equal(everything[k], v)
    got = hdf5.import_tensors(tmpfile, [r"/Kdata/KX_E\d"])        # full match: not KX_E10
    assert sorted(got) == ["/Kdata/KX_E0"]
    got = hdf5.import_tensors(tmpfile, [r"^/Kdata/KX_E.*", r".*/KY_E0"])
    assert sorted(got) == ["/Kdata/KX_E0", "/Kdata/KX_E10", "/Kdata/sub/KY_E0"]


def test_disallowed_dtype_and_missing(tmpfile):
    hdf5.export_tensors({"/i": torch.arange(5, dtype=torch.int32)}, tmpfile)
    with pytest.raises(RuntimeError, match="disallowed dtype"):
        hdf5.import_tensor(tmpfile, "/i")
    with pytest.raises(RuntimeError, match="no dataset"):
        hdf5.import_tensor(tmpfile, "/nope")
    with pytest.raises(RuntimeError, match="does not exist"):
        hdf5.import_tensor(tmpfile + ".missing", "/i")


def test_raw_kspace_loader(tmpfile):
    """The reference's layout: per encode KX/KY/KZ/KW_E{e} (any shape, flattened;
    coordinates scaled by its literal 3.141592/160) and KData_E{e}_C{c} stacked
    over c = 0, 1, ... while present."""
    rng = np.random.default_rng(3)
    E, C, shape = 2, 3, (4, 25)
    ts = {}
    for e in range(E):
        for a in "XYZW":
            ts[f"/Kdata/K{a}_E{e}"] = torch.from_numpy(rng.standard_normal(shape).astype(np.float32))
        for c in range(C):
            ts[f"/Kdata/KData_E{e}_C{c}"] = torch.from_numpy(
                (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64))
    ts["/Kdata/KData_E1_C5"] = torch.zeros(shape, dtype=torch.complex64)   # after a gap: not a coil
    ts["/Other/KX_E0"] = torch.zeros(1)
    hdf5.export_tensors(ts, tmpfile)
    enc = hdf5.import_raw_kspace(tmpfile)
    assert len(enc) == E
    for e, r in enumerate(enc):
        for a, k in zip("XYZ", r.coords):
            assert k.shape == (100,)
            assert torch.equal(k, ts[f"/Kdata/K{a}_E{e}"].flatten() * (3.141592 / 160.0))
        assert torch.equal(r.weights, ts[f"/Kdata/KW_E{e}"].flatten())
        assert r.kdata.shape == (C, 100) and r.kdata.dtype == torch.complex64
        for c in range(C):
            assert torch.equal(r.kdata[c], ts[f"/Kdata/KData_E{e}_C{c}"].flatten())
    only1 = hdf5.import_raw_kspace(tmpfile, encodes=[1])
    assert len(only1) == 1 and torch.equal(only1[0].kdata, enc[1].kdata)
    with pytest.raises(RuntimeError, match="no dataset"):
        hdf5.import_raw_kspace(tmpfile, encodes=[7])


def test_cpp_mirror_reads_and_writes_the_same_files(tmp_path):
    """include/hasty/hdf5.hpp (the C++ mirror, on the HDF5 C API) reads a file
    the Python module wrote, and writes one the Python module reads back."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    libdir = os.path.dirname(os.path.realpath(hdf5._find_library()))
    incdir = os.path.join(os.path.dirname(libdir), "include")
    if not os.path.exists(os.path.join(incdir, "hdf5.h")):
        pytest.skip("hdf5.h not found next to the HDF5 library")
    src_file, out_file = tmp_path / "in.h5", tmp_path / "out.h5"
    g = torch.Generator().manual_seed(7)
    kd = torch.randn((3, 5), generator=g, dtype=torch.complex64)
    hdf5.export_tensors({"/Kdata/KData_E0_C1": kd, "/Kdata/KX_E0": torch.arange(6.0).reshape(2, 3),
                         "/Kdata/KX_E1": torch.ones(2), "/Other/KX_E0": torch.zeros(1),
                         "/ints": torch.arange(3, dtype=torch.int32)}, str(src_file))
    src = tmp_path / "h5.cpp"
    src.write_text(r'''
#include <hasty/hdf5.hpp>
#include <cstdio>
using namespace hasty;
int main(int argc, char** argv) {
    auto k = import_tensor(argv[1], "/Kdata/KData_E0_C1").as<c64_t, 2>();
    double sr = 0, si = 0;
    for (auto v : k.vec()) { sr += v.real(); si += v.imag(); }
    std::printf("dims %lld %lld sum %.9g %.9g\n", (long long)k.shape<0>(), (long long)k.shape<1>(), sr, si);
    auto sel = import_tensors(argv[1], std::vector<std::regex>{std::regex("^/Kdata/KX_E.*")});
    auto x0 = sel.at("/Kdata/KX_E0").as<f32_t, 2>();
    std::printf("matched %zu x0 %g %g\n", sel.size(), (double)x0.vec()[1], (double)x0.vec()[5]);
    try { (void)import_tensor(argv[1], "/ints"); return 3; }
    catch (const std::runtime_error& e) { if (std::string(e.what()) != "disallowed dtype") return 4; }
    tensor<cpu_t, c128_t, 2> w({2, 3});
    for (int i = 0; i < 6; ++i) w.vec()[size_t(i)] = c128_t(0.25 * i, -1.0 - i);
    export_tensor(w, argv[2], "/grp/w");
    return 0;
}
''')
    exe = tmp_path / "h5"
    # link through a directory holding only libhdf5: the library's own
    # directory may carry an older libstdc++ that must not shadow the system one
    private = tmp_path / "h5lib"
    private.mkdir()
    real = os.path.realpath(hdf5._find_library())
    import glob
    for name in {"libhdf5.so"} | {os.path.basename(q) for q in glob.glob(os.path.join(libdir, "libhdf5.so*"))}:
        os.symlink(real, private / name)
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-D__HIP_PLATFORM_AMD__",
                        "-I", os.path.join(root, "include"), "-I", "/opt/rocm/include", "-I", incdir, str(src),
                        "-o", str(exe), "-L", str(private), "-lhdf5", "-Wl,-rpath," + str(private)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe), str(src_file), str(out_file)], capture_output=True, text=True)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    lines = r.stdout.splitlines()
    d = lines[0].split()
    assert d[1:3] == ["3", "5"]
    assert abs(float(d[4]) - kd.real.double().sum().item()) < 1e-5
    assert abs(float(d[5]) - kd.imag.double().sum().item()) < 1e-5
    assert lines[1].split() == ["matched", "2", "x0", "1", "5"]
    back = hdf5.import_tensor(str(out_file), "/grp/w")
    ref = torch.complex(0.25 * torch.arange(6, dtype=torch.float64),
                        -1.0 - torch.arange(6, dtype=torch.float64)).reshape(2, 3)
    assert back.dtype == torch.complex128 and torch.equal(back, ref)
