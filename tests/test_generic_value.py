"""CPU: the generic_value wire format (hastycucompute_amd/generic_value.py and
include/hasty/generic_value.hpp; the reference's generic_value.cppm:20-37,
261-437, 456-666).

The reference holds no serialised fixtures, so the format is pinned by:
* a byte string assembled field by field from the reference's header struct;
* the ATen ScalarType codes, read from torch's own c10 header;
* exact round trips;
* the C++ and Python sides reading each other's bytes.
"""
import io
import os
import struct
import subprocess

import pytest
import torch

from hastycucompute_amd import generic_value as gv

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_byte_layout_of_a_tensor():
    t = torch.tensor([1.0, -2.0], dtype=torch.float32)
    want = bytes.fromhex(
        "0100000000000000"                      # tag 1 (tensor) as int64
        "00" "06" "01" "ff" "00000000"          # cpu, Float, ndim 1, index -1, padding
        "0200000000000000"                      # total_elements 2
        "0200000000000000"                      # shape[0] = 2
        "0000803f" "000000c0")                  # 1.0f, -2.0f
    assert gv.serialize(t) == want
    back = gv.deserialize(want)
    assert back.dtype == torch.float32 and torch.equal(back, t)


def test_byte_layout_of_containers():
    v = {"k": [None, gv.Tuple([])]}
    want = (struct.pack("<qq", 3, 1) + struct.pack("<q", 1) + b"k"   # dict, 1 entry, key "k"
            + struct.pack("<qq", 2, 2)                               # list of 2
            + struct.pack("<q", 0)                                   # None
            + struct.pack("<qq", 4, 0))                              # empty tuple
    assert gv.serialize(v) == want
    back = gv.deserialize(want)
    assert back == {"k": [None, ()]} and isinstance(back["k"][1], gv.Tuple)


def test_scalar_type_codes_match_aten(tmp_path):
    """The codes in the header are c10::ScalarType values: read them from
    torch's own header."""
    from torch.utils.cpp_extension import include_paths
    names = {torch.uint8: "Byte", torch.int8: "Char", torch.int16: "Short", torch.int32: "Int",
             torch.int64: "Long", torch.float16: "Half", torch.float32: "Float", torch.float64: "Double",
             torch.complex32: "ComplexHalf", torch.complex64: "ComplexFloat", torch.complex128: "ComplexDouble",
             torch.bool: "Bool", torch.bfloat16: "BFloat16"}
    src = tmp_path / "sc.cpp"
    body = "".join(f'std::printf("{n} %d\\n", int(c10::ScalarType::{n}));' for n in names.values())
    src.write_text("#include <c10/core/ScalarType.h>\n#include <cstdio>\nint main(){" + body + "}\n")
    exe = tmp_path / "sc"
    r = subprocess.run(["g++", "-std=c++17", "-I", include_paths()[0], str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("torch's c10 headers do not compile here: " + r.stderr[-300:])
    out = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                       check=True).stdout.splitlines())
    for dt, n in names.items():
        assert gv._SCALAR_CODE[dt] == int(out[n]), n


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64, torch.complex64, torch.complex128, torch.int16,
                                   torch.int32, torch.int64, torch.bool, torch.uint8, torch.bfloat16])
@pytest.mark.parametrize("shape", [(5,), (2, 3, 4), (), (0, 3)])
def test_round_trip_dtypes(dtype, shape):
    g = torch.Generator().manual_seed(2)
    t = (torch.randn(shape, generator=g, dtype=torch.float64) * 10).to(dtype)
    back = gv.deserialize(gv.serialize(t))
    assert back.dtype == dtype and tuple(back.shape) == shape and torch.equal(back, t)


def test_round_trip_nested_and_streamed():
    g = torch.Generator().manual_seed(3)
    v = {"coords": [torch.rand(7, generator=g) for _ in range(3)],
         "kdata": torch.randn(4, 7, generator=g, dtype=torch.complex64),
         "pair": (torch.arange(3), {"inner": None}),
         "empty": []}
    b = gv.serialize(v)
    for back in (gv.deserialize(b), gv.deserialize(io.BytesIO(b))):
        assert list(back) == list(v)
        assert all(torch.equal(a, c) for a, c in zip(back["coords"], v["coords"]))
        assert torch.equal(back["kdata"], v["kdata"])
        assert isinstance(back["pair"], tuple) and torch.equal(back["pair"][0], v["pair"][0])
        assert back["pair"][1] == {"inner": None} and back["empty"] == []
    # a non-contiguous view is written as its contiguous data
    x = torch.arange(12.0).reshape(3, 4).t()
    assert torch.equal(gv.deserialize(gv.serialize(x)), x)


def test_errors():
    b = gv.serialize(torch.ones(3))
    with pytest.raises(RuntimeError, match="Buffer too small to deserialize tensor data"):
        gv.deserialize(b[:-1])
    with pytest.raises(RuntimeError, match="Buffer too small to deserialize type tag"):
        gv.deserialize(b[:5])
    bad = bytearray(b)
    bad[16:24] = struct.pack("<q", 4)          # total_elements 4 for shape (3,)
    with pytest.raises(RuntimeError, match="Tensor shape does not match total elements in header"):
        gv.deserialize(bytes(bad))
    bad = bytearray(b)
    bad[24:32] = struct.pack("<q", -3)         # a negative size
    with pytest.raises(RuntimeError, match="Tensor shape does not match total elements in header"):
        gv.deserialize(bytes(bad))
    with pytest.raises(RuntimeError, match="unknown generic_value type tag 9"):
        gv.deserialize(struct.pack("<q", 9))
    with pytest.raises(TypeError):
        gv.serialize({1: torch.ones(1)})
    with pytest.raises(TypeError):
        gv.serialize(3.0)


def test_type_and_metadata_strings():
    t = torch.zeros(3, 224, 224)
    assert gv.type_string(t) == "Tensor[cpu,f32,3]"
    assert gv.metadata_string(t) == "Tensor[cpu,Float32,shape=(3,224,224)]"
    lst = [torch.zeros(2, 2, dtype=torch.complex64)] * 5
    assert gv.type_string(lst) == "List[Tensor[cpu,c64,2]]"
    assert gv.metadata_string(lst) == "List[n=5,elem=Tensor[cpu,Complex64,shape=(2,2)]]"
    tup = (torch.zeros(1, dtype=torch.int64), None)
    assert gv.type_string(tup) == "Tuple[Tensor[cpu,i64,1],None]"
    assert gv.type_string({}) == "Dict[String,?]" and gv.metadata_string({}) == "Dict[n=0]"
    assert gv.type_string([]) == "List[]" and gv.metadata_string([]) == "List[n=0]"
    assert gv.metadata_string({"a": torch.zeros(2, dtype=torch.float64)}) == \
        "Dict[n=1,value_type=Tensor[cpu,Float64,shape=(2)]]"


def test_cpp_mirror_reads_and_writes_the_same_bytes(tmp_path):
    """include/hasty/generic_value.hpp re-serialises Python's bytes exactly,
    builds a value Python reads back, and gives the same descriptions."""
    g = torch.Generator().manual_seed(5)
    v = {"a": [torch.randn(2, 3, generator=g), torch.randn(4, generator=g)],
         "b": (torch.randn(3, generator=g, dtype=torch.complex64), None),
         "c": torch.arange(5, dtype=torch.int64)}        # keys in sorted order, as std::map writes them
    src_bin, out_bin = tmp_path / "in.bin", tmp_path / "out.bin"
    src_bin.write_bytes(gv.serialize(v))
    src = tmp_path / "gv.cpp"
    src.write_text(r'''
#include <hasty/generic_value.hpp>
#include <cstdio>
#include <fstream>
using namespace hasty;
int main(int argc, char** argv) {
    std::ifstream in(argv[1], std::ios::binary);
    auto v = generic_value::deserialize(in);
    std::printf("%s\n%s\n", v.type_string().c_str(), v.metadata_string().c_str());
    std::ofstream echo(argv[2], std::ios::binary);
    v.serialize(echo);
    // a value of its own, from typed host tensors
    tensor<cpu_t, c64_t, 2> k({2, 2});
    for (int i = 0; i < 4; ++i) k.vec()[size_t(i)] = c64_t(float(i), -0.5f * float(i));
    tensor<cpu_t, f64_t, 1> w({3});
    w.vec() = {0.25, 0.5, 1.0};
    generic_value::dict d;
    d["kdata"] = generic_value(wire_tensor::from(k));
    d["weights"] = generic_value::make_tuple({generic_value(wire_tensor::from(w)), generic_value()});
    std::ofstream own(std::string(argv[2]) + ".own", std::ios::binary);
    generic_value(d).serialize(own);
    // a device tensor's description (header fields set by hand: no GPU here)
    wire_tensor dev;
    dev.device_type = 1; dev.device_index = 0; dev.shape = {3, 224, 224};
    dev.bytes.resize(size_t(3 * 224 * 224 * 4));
    generic_value gd(generic_value::list{generic_value(dev)});
    std::printf("%s\n%s\n", gd.type_string().c_str(), gd.metadata_string().c_str());
    // typed access and its errors
    auto back = v.to_dict().at("c").to_tensor().as<int64_t, 1>();
    std::printf("c4 %lld\n", (long long)back.vec()[4]);
    try { (void)v.to_dict().at("c").to_tensor().as<f32_t, 1>(); return 3; }
    catch (const std::runtime_error&) {}
    try { std::vector<uint8_t> b = v.serialize(); b.resize(b.size() - 1); (void)generic_value::deserialize(b); return 4; }
    catch (const std::runtime_error& e) { std::printf("trunc %s\n", e.what()); }
    return 0;
}
''')
    exe = tmp_path / "gv"
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-D__HIP_PLATFORM_AMD__", "-I",
                        os.path.join(ROOT, "include"), "-I", "/opt/rocm/include", str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe), str(src_bin), str(out_bin)], capture_output=True, text=True)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    lines = r.stdout.splitlines()
    assert lines[0] == gv.type_string(v) and lines[1] == gv.metadata_string(v)
    assert out_bin.read_bytes() == src_bin.read_bytes()
    assert lines[2] == "List[Tensor[cuda,f32,3]]"
    assert lines[3] == "List[n=1,elem=Tensor[cuda:0,Float32,shape=(3,224,224)]]"
    assert lines[4] == "c4 4"
    assert lines[5] == "trunc Buffer too small to deserialize tensor data"
    own = gv.deserialize((tmp_path / "out.bin.own").read_bytes())
    k = torch.complex(torch.arange(4.0), -0.5 * torch.arange(4.0)).reshape(2, 2).to(torch.complex64)
    assert list(own) == ["kdata", "weights"] and torch.equal(own["kdata"], k)
    assert torch.equal(own["weights"][0], torch.tensor([0.25, 0.5, 1.0], dtype=torch.float64))
    assert own["weights"][1] is None
