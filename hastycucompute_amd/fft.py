"""NUFFT and Toeplitz operators on PyTorch-ROCm tensors (host mirror of
``hasty::fft``, ``cpp/corelib/lib/fft/{nufft,toeplitz}.cppm``).

Names, argument order and error behaviour follow the reference:

* ``nufft_opts`` / ``nufft_plan.make(opts)`` / ``setpts(coords)`` /
  ``execute(input, output)`` / ``free()`` -- ``nufft.cppm:78-310``;
* ``toeplitz_kernel(coords, kernel, nudata)`` -- ``nufft.cppm:386-441``;
* ``toeplitz_multiply(inout, kernel)`` -- ``nufft.cppm:448-479``;
* ``transform_toeplitz_kernel`` / ``toeplitz_multiplication`` with
  ``ToeplitzMultType`` / ``ToeplitzAccumulateType`` -- ``toeplitz.cppm:47-93``.

Every compute call goes through ``libhasty_hip.so``; tensors must live on the
GPU.  Errors raise ``RuntimeError`` (``HastyError``) like the reference's
``std::runtime_error``.
"""
from __future__ import annotations

import ctypes
import enum
import math
from dataclasses import dataclass, field
from typing import Optional, Sequence

import torch

from . import _lib
from ._lib import check, i32_array, i64_array, lib


class nufft_type(enum.IntEnum):
    TYPE_1 = 1
    TYPE_2 = 2
    TYPE_3 = 3
    FORWARD = 2
    BACKWARD = 1


class nufft_sign(enum.IntEnum):
    POS = 1
    NEG = -1
    DEFAULT_TYPE_1 = 1
    DEFAULT_TYPE_2 = -1


class nufft_upsamp_cuda(enum.IntEnum):
    UPSAMP_2_0 = 0
    UPSAMP_1_25 = 1
    DEFAULT = 2


class nufft_method_cuda(enum.IntEnum):
    GM_NO_SORT = 0
    GM_SORT = 1
    SM = 2
    DEFAULT = 3


class ToeplitzMultType(enum.IntEnum):
    NONE = 0
    MULT = 1
    MULT_CONJ = 2


class ToeplitzAccumulateType(enum.IntEnum):
    NONE = 0
    ACCUMULATE = 1


def _stream(device: torch.device | None = None) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def _ptr(t: Optional[torch.Tensor]) -> Optional[int]:
    return None if t is None else t.data_ptr()


def _require_gpu(t: torch.Tensor, name: str) -> None:
    if not isinstance(t, torch.Tensor):
        raise RuntimeError(f"{name} must be a torch.Tensor")
    if t.device.type != "cuda":
        raise RuntimeError(f"{name} must be a GPU tensor (got {t.device})")
    if not t.is_contiguous():
        raise RuntimeError(f"{name} must be contiguous")


def verify_coords(coords: Sequence[torch.Tensor]) -> None:
    """``nufft.cppm:23-39``: every coordinate vector has the same length."""
    lens = {int(c.shape[0]) for c in coords}
    if len(lens) != 1:
        if len(coords) == 2:
            raise RuntimeError("x and y coords have different lengths")
        raise RuntimeError("x,y  x,z  or  y,z  coords have different lengths")


@dataclass
class nufft_opts:
    """``nufft_opts<cuda_t,TT,DIM>`` (``nufft.cppm:78-95``) plus tiling knobs."""
    nmodes: Sequence[int]
    ntransf: int = 1
    tol: float = 1e-5
    sign: int = 0                       # 0 -> the type's default sign
    upsamp: nufft_upsamp_cuda = nufft_upsamp_cuda.DEFAULT
    method: nufft_method_cuda = nufft_method_cuda.DEFAULT
    pstream: object = None              # unused, as in the reference
    device_idx: int = 0
    tile: Sequence[int] = field(default_factory=lambda: (0, 0, 0))
    coil_group: int = 0
    max_batch: int = 0
    max_subproblem: int = 0

    def get_sign(self) -> int:
        return int(self.sign)


class nufft_plan:
    """``nufft_plan<cuda_t,TT,DIM,NT>`` over the gfx950 library.

    ``dtype`` is the real type of the plan (``torch.float32`` -> c64 data,
    ``torch.float64`` -> c128 data), the ``TT`` template parameter.
    """

    def __init__(self):
        self._h = None

    @staticmethod
    def make(opts: nufft_opts, nt: nufft_type, dtype: torch.dtype = torch.float32) -> "nufft_plan":
        nt = nufft_type(nt)
        if nt == nufft_type.TYPE_3:
            raise RuntimeError("nufft type 3 is not supported")
        if dtype not in (torch.float32, torch.float64):
            raise RuntimeError("plan dtype must be float32 or float64")
        dim = len(opts.nmodes)
        if dim not in (2, 3):
            raise RuntimeError("nufft dim must be 2 or 3")
        if opts.upsamp not in (nufft_upsamp_cuda.UPSAMP_2_0, nufft_upsamp_cuda.UPSAMP_1_25,
                               nufft_upsamp_cuda.DEFAULT):
            raise RuntimeError("nufft_make_plan encountered invalid nufft_upsamp argument")
        if opts.method not in tuple(nufft_method_cuda):
            raise RuntimeError("Invalid nufft_cuda_method")
        sign = opts.get_sign() or (1 if nt == nufft_type.TYPE_1 else -1)
        o = _lib.NufftOpts()
        o.upsamp = 1.25 if opts.upsamp == nufft_upsamp_cuda.UPSAMP_1_25 else 2.0
        o.method = int(opts.method)
        for i, t in enumerate(list(opts.tile)[:3]):
            o.tile[i] = int(t)
        o.coil_group = int(opts.coil_group)
        o.max_batch = int(opts.max_batch)
        o.max_subproblem = int(opts.max_subproblem)
        p = nufft_plan()
        p._opts = opts
        p._type = nt
        p._dim = dim
        p._dtype = dtype
        p._cdtype = torch.complex64 if dtype == torch.float32 else torch.complex128
        p._device = torch.device("cuda", torch.cuda.current_device())
        h = ctypes.c_void_p()
        with torch.cuda.device(p._device):
            check(lib().hasty_nufft_makeplan(int(nt), dim, i64_array(opts.nmodes), int(sign),
                                             int(opts.ntransf), float(opts.tol),
                                             0 if dtype == torch.float32 else 1, ctypes.byref(o),
                                             _stream(p._device), ctypes.byref(h)))
        p._h = h
        p._M = None
        p._coords = None
        return p

    def _handle(self):
        if self._h is None or not self._h.value:
            raise RuntimeError("nufft plan has been freed")
        return self._h

    def setpts(self, coords: Sequence[torch.Tensor]) -> None:
        if len(coords) != self._dim:
            raise RuntimeError(f"expected {self._dim} coordinate vectors")
        verify_coords(coords)
        for i, c in enumerate(coords):
            _require_gpu(c, f"coords[{i}]")
            if c.dtype != self._dtype:
                raise RuntimeError(f"coords[{i}] must have dtype {self._dtype}")
        M = int(coords[0].shape[0])
        ptrs = [c.data_ptr() for c in coords] + [None] * (3 - self._dim)
        with torch.cuda.device(self._device):
            check(lib().hasty_nufft_set_stream(self._handle(), _stream(self._device)))
            check(lib().hasty_nufft_setpts(self._handle(), M, *ptrs))
        self._M = M
        self._coords = list(coords)   # kept alive like the reference's borrowed pointers

    def execute(self, input: torch.Tensor, output: torch.Tensor) -> None:
        if self._M is None:
            raise RuntimeError("cufinufft: execute failed with error code: setpts not called")
        nt = int(self._opts.ntransf)
        for name, t in (("input", input), ("output", output)):
            _require_gpu(t, name)
            if t.dtype != self._cdtype:
                raise RuntimeError(f"{name} must have dtype {self._cdtype}")
        modes = tuple(int(m) for m in self._opts.nmodes)
        if self._type == nufft_type.TYPE_1:
            c, f = input, output
        else:
            f, c = input, output
        if c.dim() != 2 or c.shape[0] != nt or c.shape[1] != self._M:
            if self._type == nufft_type.TYPE_1 and c.dim() == 2 and c.shape[0] != nt:
                raise RuntimeError("input first dimension must match ntransf")
            raise RuntimeError(f"nonuniform data must have shape [{nt}, {self._M}]")
        if tuple(f.shape) != (nt,) + modes:
            raise RuntimeError(f"mode data must have shape {(nt,) + modes}")
        with torch.cuda.device(self._device):
            check(lib().hasty_nufft_set_stream(self._handle(), _stream(self._device)))
            check(lib().hasty_nufft_execute(self._handle(), c.data_ptr(), f.data_ptr()))

    def coord_grad(self, h: torch.Tensor, w: torch.Tensor) -> list:
        """Type-2 plans: ``out[a][j] = Re sum_t conj(w[t, j]) d/dk_{j,a} type2(h)[t, j]``
        for the coordinates of the last ``setpts`` (one real tensor [M] per
        coordinate, in the caller's point order), by interpolation with the
        derivative of the window.  No counterpart in the reference."""
        if self._type != nufft_type.TYPE_2:
            raise RuntimeError("the coordinate gradient runs on a type-2 plan")
        if self._M is None:
            raise RuntimeError("setpts not called")
        nt = int(self._opts.ntransf)
        modes = tuple(int(m) for m in self._opts.nmodes)
        for name, t in (("h", h), ("w", w)):
            _require_gpu(t, name)
            if t.dtype != self._cdtype:
                raise RuntimeError(f"{name} must have dtype {self._cdtype}")
        if tuple(h.shape) != (nt,) + modes:
            raise RuntimeError(f"h must have shape {(nt,) + modes}")
        if tuple(w.shape) != (nt, self._M):
            raise RuntimeError(f"w must have shape [{nt}, {self._M}]")
        out = torch.empty((self._dim, self._M), dtype=self._dtype, device=self._device)
        with torch.cuda.device(self._device):
            check(lib().hasty_nufft_set_stream(self._handle(), _stream(self._device)))
            check(lib().hasty_nufft_coord_grad(self._handle(), h.data_ptr(), w.data_ptr(), out.data_ptr()))
        return [out[a] for a in range(self._dim)]

    def info(self) -> dict:
        w = ctypes.c_int()
        beta = ctypes.c_double()
        nf = (ctypes.c_int64 * 3)()
        tile = (ctypes.c_int32 * 3)()
        nbytes = ctypes.c_int64()
        check(lib().hasty_nufft_info(self._handle(), ctypes.byref(w), ctypes.byref(beta), nf, tile,
                                     ctypes.byref(nbytes)))
        return {"width": w.value, "beta": beta.value, "nf": tuple(nf), "tile": tuple(tile),
                "device_bytes": nbytes.value}

    def nmodes(self):
        return tuple(self._opts.nmodes)

    def ntransf(self) -> int:
        return int(self._opts.ntransf)

    def tol(self) -> float:
        return float(self._opts.tol)

    def free(self) -> None:
        if self._h is not None and self._h.value:
            check(lib().hasty_nufft_destroy(self._h))
        self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def nufft_backward_cpu_over_cuda(dimshape: Sequence[int], input: torch.Tensor,
                                 coords: Sequence[torch.Tensor], tol: float = 1e-6,
                                 batch: int = 0) -> torch.Tensor:
    """``nufft.cppm:316-382``: adjoint NUFFT of host data ``[C, M]`` with GPU
    coordinates, result on the host ``[C, N...]``.  All coils (or ``batch``
    coils at a time) go through one plan execute instead of one per coil."""
    device = coords[0].device
    C = int(input.shape[0])
    nb = C if batch <= 0 else min(batch, C)
    rdt = coords[0].dtype
    cdt = torch.complex64 if rdt == torch.float32 else torch.complex128
    plan = nufft_plan.make(nufft_opts(nmodes=tuple(dimshape), ntransf=nb, tol=tol,
                                      sign=nufft_sign.DEFAULT_TYPE_1), nufft_type.TYPE_1, rdt)
    plan.setpts(coords)
    out = torch.empty((C,) + tuple(dimshape), dtype=input.dtype)
    for c0 in range(0, C, nb):
        chunk = input[c0:c0 + nb]
        if chunk.shape[0] < nb:
            pad = torch.zeros((nb,) + tuple(chunk.shape[1:]), dtype=chunk.dtype)
            pad[:chunk.shape[0]] = chunk
            chunk = pad
        g_in = chunk.to(device=device, dtype=cdt).contiguous()
        g_out = torch.empty((nb,) + tuple(dimshape), dtype=cdt, device=device)
        plan.execute(g_in, g_out)
        n = min(nb, C - c0)
        out[c0:c0 + n] = g_out[:n].to("cpu", dtype=input.dtype)
    plan.free()
    return out


def toeplitz_kernel(coords: Sequence[torch.Tensor], kernel: torch.Tensor,
                    nudata: Optional[torch.Tensor] = None, tol: float = 1e-9) -> None:
    """``nufft.cppm:386-441``: fill ``kernel [2N...]`` (c64) with the Toeplitz
    kernel of the trajectory, weights ``nudata [1, M]`` (None -> ones)."""
    verify_coords(coords)
    _require_gpu(kernel, "kernel")
    if kernel.dtype != torch.complex64:
        raise RuntimeError("kernel dtype must be complex float")
    dim = kernel.dim()
    if len(coords) != dim:
        raise RuntimeError("kernel rank must equal the number of coordinate vectors")
    for i, c in enumerate(coords):
        _require_gpu(c, f"coords[{i}]")
        if c.dtype != torch.float32:
            raise RuntimeError("coords must be float32")
    n = [int(s) // 2 for s in kernel.shape]
    if any(2 * m != s for m, s in zip(n, kernel.shape)):
        raise RuntimeError("kernel dimensions must be even (2N)")
    M = int(coords[0].shape[0])
    w = None
    if nudata is not None:
        _require_gpu(nudata, "nudata")
        if nudata.dtype != torch.complex64:
            raise RuntimeError("nudata must be complex64")
        if nudata.shape[-1] != M:
            raise RuntimeError("shape<1>() of nudata didn't match length of coord vectors")
        w = nudata.data_ptr()
    ptrs = [c.data_ptr() for c in coords] + [None] * (3 - dim)
    check(lib().hasty_toeplitz_kernel(dim, i64_array(n), M, *ptrs, w, kernel.data_ptr(), float(tol),
                                      _stream(kernel.device)))


def _toeplitz_dims(x: torch.Tensor, kernel: torch.Tensor):
    """Shape checks of ``toeplitz.cppm:630-640,681-692``.  The 2N grid of
    every axis must factor as 2^a 3^b 5^c (the Stockham radices)."""
    d = kernel.dim()
    if x.dim() != d + 1:
        raise RuntimeError("batched tensor must have rank kernel.dim() + 1")
    n = [int(s) for s in x.shape[1:]]
    if tuple(kernel.shape) != tuple(2 * m for m in n):
        raise RuntimeError("kernel must have shape 2N in every dimension")
    return d, n


def toeplitz_multiply(inout: torch.Tensor, kernel: torch.Tensor) -> None:
    """``nufft.cppm:448-479`` in place on ``inout [B, N...]``."""
    _require_gpu(inout, "inout")
    _require_gpu(kernel, "kernel")
    for name, t in (("inout", inout), ("kernel", kernel)):
        if t.dtype != torch.complex64:
            raise RuntimeError(f"{name} dtype must be complex float")
    d, n = _toeplitz_dims(inout, kernel)
    check(lib().hasty_toeplitz_multiply(d, i64_array(n), int(inout.shape[0]), inout.data_ptr(),
                                        kernel.data_ptr(), _stream(inout.device)))


def transform_toeplitz_kernel(kernel: torch.Tensor, clear_vkfft_plan: bool = False) -> None:
    """``toeplitz.cppm:510-609``.  The kernel stays in natural frequency order in
    this library, so there is nothing to re-lay-out; shape is validated."""
    _require_gpu(kernel, "kernel")
    n = [int(s) // 2 for s in kernel.shape]
    check(lib().hasty_transform_toeplitz_kernel(kernel.dim(), i64_array(n), kernel.data_ptr(),
                                                _stream(kernel.device)))


def toeplitz_multiplication(input: torch.Tensor, output: torch.Tensor, kernel: torch.Tensor,
                            scratch: Optional[torch.Tensor], mult1: Optional[torch.Tensor],
                            mult2: Optional[torch.Tensor],
                            input_output_mult_type: ToeplitzMultType,
                            input_mult1_type: ToeplitzMultType, output_mult1_type: ToeplitzMultType,
                            input_mult2_type: ToeplitzMultType, output_mult2_type: ToeplitzMultType,
                            accumulate_type: ToeplitzAccumulateType) -> None:
    """``toeplitz.cppm:611-712``: batched Toeplitz product with diagonal operands."""
    for name, t in (("input", input), ("output", output), ("kernel", kernel)):
        _require_gpu(t, name)
        if t.dtype != torch.complex64:
            raise RuntimeError(f"{name} dtype must be complex float")
    if input.shape != output.shape:
        raise RuntimeError("input and output must have the same size")
    if input.device != output.device:
        raise RuntimeError("input and output must be on the same device")
    if input.device != kernel.device:
        raise RuntimeError("input and kernel must be on the same device")
    d, n = _toeplitz_dims(input, kernel)
    for name, m in (("mult1", mult1), ("mult2", mult2)):
        if m is not None:
            _require_gpu(m, name)
            if m.dtype != torch.complex64 or list(m.shape) != n:
                raise RuntimeError(f"{name} must be complex float of shape {n}")
    try:
        types = [int(ToeplitzMultType(t)) for t in (input_output_mult_type, input_mult1_type,
                                                     output_mult1_type, input_mult2_type,
                                                     output_mult2_type)]
        acc = int(ToeplitzAccumulateType(accumulate_type))
    except ValueError as e:   # the reference throws std::runtime_error (toeplitz.cppm:794-840)
        raise RuntimeError(f"invalid Toeplitz enum value: {e}") from None
    B = int(input.shape[0])
    sbytes = 0
    sptr = None
    if scratch is not None:
        _require_gpu(scratch, "scratch")
        sbytes = scratch.numel() * scratch.element_size()
        sptr = scratch.data_ptr()
    check(lib().hasty_toeplitz_multiplication(d, i64_array(n), B, input.data_ptr(), output.data_ptr(),
                                              kernel.data_ptr(), sptr, sbytes, _ptr(mult1), _ptr(mult2),
                                              i32_array(types), acc, _stream(input.device)))


def toeplitz_scratch_bytes(shape: Sequence[int], batch: int) -> int:
    r = lib().hasty_toeplitz_scratch_bytes(len(shape), i64_array(shape), int(batch))
    if r < 0:
        check(_lib.HASTY_OK + 1)
    return int(r)


def kb_width(tol: float, upsamp: float = 2.0, dim: int = 3, dtype: torch.dtype = torch.float32) -> int:
    """Kernel width a plan of real type ``dtype`` uses for a tolerance:
    ``ceil(-log10(tol / 10))`` at upsampling 2, else the smallest width whose
    estimated error (window aliasing plus the rounding floor of the plan's
    arithmetic) is <= tol; 17 when none up to 16 is (``kb_width_for`` in
    ``csrc/kb.hpp``).  A live plan reports its width in ``nufft_plan.info()``."""
    w = ctypes.c_int()
    check(lib().hasty_nufft_kernel_width(float(tol), float(upsamp), int(dim),
                                         1 if dtype == torch.float64 else 0, ctypes.byref(w)))
    return int(w.value)
