"""The generic_value wire format: tensors, lists, tuples and string-keyed
dicts of them, serialised to bytes (the reference's
``cpp/corelib/lib/generic_value/generic_value.cppm:20-37, 261-437, 456-666``).

Layout (all integers little-endian):

* a value opens with its type tag as an int64: 0 None, 1 tensor, 2 list,
  3 dict, 4 tuple (``value_type``, :20-26);
* tensor: a 16-byte header ``{u8 device_type (0 cpu, 1 cuda), u8 scalar_type
  (the ATen ScalarType code), u8 ndim, i8 device_index (-1 on the host),
  4 bytes of padding (written as zeros), i64 total_elements}`` (:30-37, the
  C struct's natural alignment), then ``ndim`` int64 sizes, then the
  contiguous element bytes;
* list / tuple: int64 count, then each element;
* dict: int64 count, then per entry an int64 key length, the key's bytes and
  the value.

The tag width follows the reference's writer (:349, an 8-byte integer); its
reader takes a single byte (:400), which would read the writer's output out of
step, so the reader here takes the 8 bytes the writer wrote.  Device tensors
are written from a host copy and read back onto ``cuda:<device_index>``
unless ``device`` is given.

``type_string`` and ``metadata_string`` give the reference's descriptions
(:261-337): ``Tensor[cpu,f32,3]``, ``List[Tensor[cuda,c64,2]]``,
``Tensor[cuda:0,Float32,shape=(3,224,224)]``, ``List[n=5,elem=...]``.

Host-side format code: no device kernels are involved.
"""
import struct

import numpy as np
import torch

NONE, TENSOR, VECTOR, DICT, TUPLE = 0, 1, 2, 3, 4

# ATen ScalarType codes (c10/core/ScalarType.h) for the dtypes the format names
_SCALAR_CODE = {
    torch.uint8: 0, torch.int8: 1, torch.int16: 2, torch.int32: 3, torch.int64: 4, torch.float16: 5,
    torch.float32: 6, torch.float64: 7, torch.complex32: 8, torch.complex64: 9, torch.complex128: 10,
    torch.bool: 11, torch.bfloat16: 15,
}
_CODE_SCALAR = {v: k for k, v in _SCALAR_CODE.items()}
# generic_value.cppm:575-605
_TYPENAME = {torch.float32: "f32", torch.float64: "f64", torch.complex64: "c64", torch.complex128: "c128",
             torch.int16: "i16", torch.int32: "i32", torch.int64: "i64", torch.bool: "b8", torch.uint8: "u8"}
_LONGNAME = {torch.float32: "Float32", torch.float64: "Float64", torch.complex64: "Complex64",
             torch.complex128: "Complex128", torch.int32: "Int32", torch.int64: "Int64", torch.int16: "Int16",
             torch.uint8: "UInt8", torch.bool: "Bool"}

_HEADER = struct.Struct("<BBBb4xq")   # 16 bytes
_I64 = struct.Struct("<q")


class Tuple(tuple):
    """Marks a tuple value (tag 4); a plain tuple is written as one too."""


def _tag(v):
    if v is None:
        return NONE
    if isinstance(v, torch.Tensor):
        return TENSOR
    if isinstance(v, tuple):
        return TUPLE
    if isinstance(v, list):
        return VECTOR
    if isinstance(v, dict):
        return DICT
    raise TypeError(f"generic_value holds tensors, lists, tuples and dicts of them, not {type(v).__name__}")


def _write(v, out):
    t = _tag(v)
    out.append(_I64.pack(t))
    if t == TENSOR:
        if v.dtype not in _SCALAR_CODE:
            raise TypeError(f"dtype {v.dtype} has no ScalarType code in the format")
        cuda = v.device.type == "cuda"
        out.append(_HEADER.pack(1 if cuda else 0, _SCALAR_CODE[v.dtype], v.dim(),
                                (v.device.index or 0) if cuda else -1, v.numel()))
        out.append(struct.pack(f"<{v.dim()}q", *v.shape))
        host = v.detach().to("cpu").contiguous()
        if host.numel():
            out.append(host.reshape(-1).view(torch.uint8).numpy().tobytes())
    elif t in (VECTOR, TUPLE):
        out.append(_I64.pack(len(v)))
        for e in v:
            _write(e, out)
    elif t == DICT:
        out.append(_I64.pack(len(v)))
        for k, e in v.items():
            if not isinstance(k, str):
                raise TypeError("generic_value dict keys are strings")
            kb = k.encode()
            out.append(_I64.pack(len(kb)))
            out.append(kb)
            _write(e, out)


def serialize(value) -> bytes:
    """The bytes of ``value`` (None, a tensor, or lists / tuples / str-keyed
    dicts of these, nested)."""
    out = []
    _write(value, out)
    return b"".join(out)


class _Reader:
    def __init__(self, src):
        self.mv = memoryview(src) if isinstance(src, (bytes, bytearray, memoryview)) else None
        self.f = None if self.mv is not None else src
        self.pos = 0

    def take(self, n, what):
        if n < 0:
            raise RuntimeError(f"negative length for {what}")
        if self.mv is not None:
            if self.pos + n > len(self.mv):
                raise RuntimeError(f"Buffer too small to deserialize {what}")
            b = self.mv[self.pos:self.pos + n]
            self.pos += n
            return b
        b = self.f.read(n)
        if len(b) != n:
            raise RuntimeError(f"Buffer too small to deserialize {what}")
        return memoryview(b)

    def i64(self, what):
        return _I64.unpack(self.take(8, what))[0]


def _read(r, device):
    t = r.i64("type tag")
    if t == NONE:
        return None
    if t == TENSOR:
        dev_type, code, ndim, dev_idx, total = _HEADER.unpack(r.take(_HEADER.size, "tensor header"))
        shape = list(struct.unpack(f"<{ndim}q", r.take(8 * ndim, "shape")))
        if any(d < 0 for d in shape) or total < 0 or int(np.prod(shape, dtype=object)) != total:
            raise RuntimeError("Tensor shape does not match total elements in header")
        if code not in _CODE_SCALAR:
            raise RuntimeError(f"unknown scalar type code {code}")
        dtype = _CODE_SCALAR[code]
        nbytes = total * torch.empty((), dtype=dtype).element_size()
        raw = bytearray(r.take(nbytes, "tensor data"))
        flat = torch.frombuffer(raw, dtype=torch.uint8) if nbytes else torch.empty(0, dtype=torch.uint8)
        x = flat.view(dtype).reshape(shape)
        target = device if device is not None else (f"cuda:{dev_idx}" if dev_type == 1 else "cpu")
        return x.to(target)
    if t in (VECTOR, TUPLE):
        n = r.i64("length")
        items = [_read(r, device) for _ in range(n)]
        return Tuple(items) if t == TUPLE else items
    if t == DICT:
        n = r.i64("length")
        d = {}
        for _ in range(n):
            k = bytes(r.take(r.i64("key length"), "key")).decode()
            d[k] = _read(r, device)
        return d
    raise RuntimeError(f"unknown generic_value type tag {t}")


def deserialize(src, device=None):
    """Reads one value from bytes (or an object with ``read(n)``).  Tensors
    go to the device recorded in their header unless ``device`` is given."""
    return _read(_Reader(src), device)


def type_string(v) -> str:
    """generic_value.cppm:261-290 (lists and dicts described by their first
    element)."""
    t = _tag(v)
    if t == NONE:
        return "None"
    if t == TENSOR:
        dev = "cuda" if v.device.type == "cuda" else "cpu"
        return f"Tensor[{dev},{_TYPENAME.get(v.dtype, 'unknown')},{v.dim()}]"
    if t == VECTOR:
        return f"List[{type_string(v[0])}]" if v else "List[]"
    if t == TUPLE:
        return "Tuple[" + ",".join(type_string(e) for e in v) + "]"
    if not v:
        return "Dict[String,?]"
    return f"Dict[String,{type_string(next(iter(v.values())))}]"


def metadata_string(v) -> str:
    """generic_value.cppm:302-337."""
    t = _tag(v)
    if t == NONE:
        return "None"
    if t == TENSOR:
        dev = f"cuda:{v.device.index or 0}" if v.device.type == "cuda" else "cpu"
        shape = ",".join(str(s) for s in v.shape)
        return f"Tensor[{dev},{_LONGNAME.get(v.dtype, 'Unknown')},shape=({shape})]"
    if t == VECTOR:
        return f"List[n={len(v)}" + (f",elem={metadata_string(v[0])}" if v else "") + "]"
    if t == TUPLE:
        return "Tuple[" + ",".join(metadata_string(e) for e in v) + "]"
    return f"Dict[n={len(v)}" + (f",value_type={metadata_string(next(iter(v.values())))}" if v else "") + "]"
