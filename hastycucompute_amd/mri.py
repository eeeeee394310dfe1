"""SENSE operators: host mirror of the reference's normal-operator family
(``NORMAL_T_T1_OP`` ... ``NORMAL_IDTW_T2_OP``, ``cpp/corelib/lib/mri/normal.cppm``)
and of the time-segmented SENSE forward / adjoint
(``cpp/corelib/lib/mri/{forward,adjoint}.cppm``)."""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import torch

from . import fft as hfft
from ._lib import check, i64_array, lib


class NORMAL_T_T1_OP:
    """``normal.cppm:43-93``: x -> sum_c conj(S_c) . T(S_c . x).

    ``kernel`` is the Toeplitz kernel ``[2N...]`` (``toeplitz_kernel``) and
    ``stacked_diags`` the coil maps ``[C, N...]``, both c64 on the GPU and kept
    resident (the reference re-uploads the kernel per apply,
    ``normal.cppm:75``).  All coils are processed per launch and the coil sum
    is fused into the last pass; ``coil_batch`` caps the coils per launch (the
    workspace is ``coil_batch * (2 + 4) N^d`` complex values in 3-D).
    """

    class_name = "NORMAL_T_T1_OP"

    def __init__(self, kernel: torch.Tensor, stacked_diags: torch.Tensor, coil_batch: int = 0):
        for name, t in (("kernel", kernel), ("stacked_diags", stacked_diags)):
            hfft._require_gpu(t, name)
            if t.dtype != torch.complex64:
                raise RuntimeError(f"{name} must be complex64")
        d = kernel.dim()
        if d not in (2, 3) or stacked_diags.dim() != d + 1:
            raise RuntimeError("NORMAL_T_T1_OP requires DIM in {2,3} and maps of rank DIM+1")
        n = [int(s) for s in stacked_diags.shape[1:]]
        if tuple(kernel.shape) != tuple(2 * m for m in n):
            raise RuntimeError("kernel must have shape 2N")
        self._kernel = kernel
        self._smaps = stacked_diags
        self._n = n
        self._device = kernel.device
        h = ctypes.c_void_p()
        check(lib().hasty_normal_create(d, i64_array(n), int(stacked_diags.shape[0]),
                                        stacked_diags.data_ptr(), kernel.data_ptr(), int(coil_batch),
                                        hfft._stream(self._device), ctypes.byref(h)))
        self._h = h

    @property
    def shape(self):
        return tuple(self._n)

    def apply(self, x: torch.Tensor, out: torch.Tensor) -> torch.Tensor:
        hfft._require_gpu(x, "x")
        hfft._require_gpu(out, "out")
        if list(x.shape) != self._n or x.dtype != torch.complex64:
            raise RuntimeError(f"x must be complex64 of shape {self._n}")
        if out.shape != x.shape or out.dtype != x.dtype:
            raise RuntimeError("out must match x")
        check(lib().hasty_normal_set_stream(self._h, hfft._stream(self._device)))
        check(lib().hasty_normal_apply(self._h, x.data_ptr(), out.data_ptr()))
        return out

    def __call__(self, x: torch.Tensor) -> torch.Tensor:
        return self.apply(x, torch.empty_like(x))

    def free(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            check(lib().hasty_normal_destroy(self._h))
        self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _c64(t: torch.Tensor, name: str, shape=None) -> None:
    hfft._require_gpu(t, name)
    if t.dtype != torch.complex64:
        raise RuntimeError(f"{name} must be complex64")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise RuntimeError(f"{name} must have shape {tuple(shape)}, got {tuple(t.shape)}")


class _NormalFamily:
    """A ``hasty_normal_op`` handle of the multi-kernel family (kind 1: image ->
    image with a fused coil sum; kind 2: stacked images -> stacked images).
    The device operands are referenced, not copied, and kept alive here."""

    class_name = ""

    def __init__(self, kernels, kerneldiags, weights, chunk):
        self._h = None
        if kernels.dim() not in (3, 4):
            raise RuntimeError(f"{self.class_name}: kernels must be [L, 2N...] with DIM in {{2,3}}")
        _c64(kernels, "kernels")
        self._n = [int(s) // 2 for s in kernels.shape[1:]]
        self._d = len(self._n)
        if tuple(kernels.shape[1:]) != tuple(2 * m for m in self._n):
            raise RuntimeError("kernels must have even extents 2N")
        self._L = int(kernels.shape[0])
        if kerneldiags is not None:
            _c64(kerneldiags, "kerneldiags")
            if tuple(kerneldiags.shape[1:]) != tuple(self._n):
                raise RuntimeError(f"kerneldiags must have shape [L, {self._n}]")
            # the reference checks this when the operator is applied (normal.cppm:222-224)
            self._mismatch = int(kerneldiags.shape[0]) != self._L
        else:
            self._mismatch = False
        self._kernels = kernels
        self._kerneldiags = kerneldiags
        self._weights = weights
        self._chunk = int(chunk)
        self._device = kernels.device

    def _check_call(self):
        if self._mismatch:
            raise RuntimeError("Number of kernels and kernel diagonals must match.")

    def _apply(self, x: torch.Tensor, out: torch.Tensor) -> torch.Tensor:
        check(lib().hasty_normal_set_stream(self._h, hfft._stream(self._device)))
        check(lib().hasty_normal_apply(self._h, x.data_ptr(), out.data_ptr()))
        return out

    def free(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            check(lib().hasty_normal_destroy(self._h))
        self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class _NormalT1(_NormalFamily):
    """x -> sum_c conj(S'_c) . sum_l conj(R_l) . T_l(R_l . S_c . x)."""

    def __init__(self, kernels, kerneldiags, stacked_diags, weights=None, coil_batch=0):
        super().__init__(kernels, kerneldiags, weights, coil_batch)
        _c64(stacked_diags, "stacked_diags")
        if tuple(stacked_diags.shape[1:]) != tuple(self._n):
            raise RuntimeError(f"stacked_diags must have shape [C, {self._n}]")
        C = int(stacked_diags.shape[0])
        if weights is not None:
            _c64(weights, "weights", (C, C))
        self._smaps = stacked_diags
        if self._mismatch:
            return
        h = ctypes.c_void_p()
        check(lib().hasty_normal_idt_create(self._d, i64_array(self._n), C, stacked_diags.data_ptr(),
                                            self._L, kernels.data_ptr(), hfft._ptr(kerneldiags),
                                            hfft._ptr(weights), self._chunk,
                                            hfft._stream(self._device), ctypes.byref(h)))
        self._h = h

    def apply(self, x: torch.Tensor, out: torch.Tensor) -> torch.Tensor:
        self._check_call()
        _c64(x, "x", self._n)
        _c64(out, "out", self._n)
        return self._apply(x, out)

    def __call__(self, x: torch.Tensor) -> torch.Tensor:
        return self.apply(x, torch.empty_like(x))


class _NormalT2(_NormalFamily):
    """Psi_i -> sum_j W_ij conj(X) . sum_l conj(R_l) . T_l(R_l . X . Psi_j)."""

    def __init__(self, kernels, kerneldiags, x, weights=None, stack_batch=0):
        super().__init__(kernels, kerneldiags, weights, stack_batch)
        _c64(x, "x", self._n)
        if weights is not None:
            _c64(weights, "weights")
            if weights.dim() != 2 or weights.shape[0] != weights.shape[1]:
                raise RuntimeError("weights must be a square [B, B] matrix")
        self._x = x
        self._B = None

    def __call__(self, stacked_diags: torch.Tensor) -> torch.Tensor:
        self._check_call()
        _c64(stacked_diags, "stacked_diags")
        if tuple(stacked_diags.shape[1:]) != tuple(self._n):
            raise RuntimeError(f"stacked_diags must have shape [B, {self._n}]")
        B = int(stacked_diags.shape[0])
        if self._weights is not None and B != int(self._weights.shape[0]):
            raise RuntimeError("the number of stacked diagonals must match the weights")
        if B != self._B:
            self.free()
            h = ctypes.c_void_p()
            check(lib().hasty_normal_t2_create(self._d, i64_array(self._n), B, self._x.data_ptr(), self._L,
                                               self._kernels.data_ptr(), hfft._ptr(self._kerneldiags),
                                               hfft._ptr(self._weights), self._chunk,
                                               hfft._stream(self._device), ctypes.byref(h)))
            self._h = h
            self._B = B
        return self._apply(stacked_diags, torch.empty_like(stacked_diags))


class NORMAL_T_T2_OP(_NormalT2):
    """``normal.cppm:113-160``: (B Psi)_i = conj(x) . T(x . Psi_i)."""

    class_name = "NORMAL_T_T2_OP"

    def __init__(self, kernel: torch.Tensor, x: torch.Tensor, stack_batch: int = 0):
        super().__init__(kernel.unsqueeze(0), None, x, None, stack_batch)


class NORMAL_IDT_T1_OP(_NormalT1):
    """``normal.cppm:185-272``: x -> sum_l R_l^H [sum_c S_c^H T_l S_c] R_l x,
    i.e. sum_c conj(S_c) . sum_l conj(R_l) . T_l(R_l . S_c . x); kernels
    ``[L, 2N...]``, kerneldiags ``[L, N...]``, stacked_diags ``[C, N...]``."""

    class_name = "NORMAL_IDT_T1_OP"

    def __init__(self, kernels, kerneldiags, stacked_diags, store_module: bool = False, coil_batch: int = 0):
        super().__init__(kernels, kerneldiags, stacked_diags, None, coil_batch)


class NORMAL_IDT_T2_OP(_NormalT2):
    """``normal.cppm:292-375``: (B Psi)_i = conj(x) . sum_l conj(R_l) . T_l(R_l . x . Psi_i)."""

    class_name = "NORMAL_IDT_T2_OP"

    def __init__(self, kernels, kerneldiags, x, stack_batch: int = 0):
        super().__init__(kernels, kerneldiags, x, None, stack_batch)


class NORMAL_IDTW_T1_OP(_NormalT1):
    """``normal.cppm:405-564``: x -> sum_{i,j} W_ij conj(S_i) . sum_l conj(R_l) . T_l(R_l . S_j . x).

    The coil sum stays fused in the last FFT pass: the maps on the output
    side are S'_j = sum_i conj(W_ij) S_i, formed once at construction."""

    class_name = "NORMAL_IDTW_T1_OP"

    def __init__(self, kernels, kerneldiags, stacked_diags, weights, store_module: bool = False,
                 coil_batch: int = 0):
        super().__init__(kernels, kerneldiags, stacked_diags, weights, coil_batch)


class NORMAL_IDTW_T2_OP(_NormalT2):
    """``normal.cppm:586-735``: (B Psi)_i = sum_j W_ij conj(x) . sum_l conj(R_l) . T_l(R_l . x . Psi_j)."""

    class_name = "NORMAL_IDTW_T2_OP"

    def __init__(self, kernels, kerneldiags, x, weights, store_module: bool = False, stack_batch: int = 0):
        super().__init__(kernels, kerneldiags, x, weights, stack_batch)


class SenseOffresonance:
    """Time-segmented SENSE forward / adjoint with all coils of a chunk batched
    through one type-2 and one type-1 plan (``hasty_sense_forward`` /
    ``hasty_sense_adjoint``):

      forward: y_c = sum_l bil_l . A(cjl_l . S_c . x)                  [C, M]
      adjoint: x_c = conj(S_c) . sum_l conj(cjl_l) . A^H(conj(bil_l) . y_c)
               [C, N...], or their coil sum with ``combine=True``.

    The reference runs one single-transform plan per coil and segment
    (``forward.cppm:36-165``, ``adjoint.cppm:33-171``); here the products
    with S, cjl and bil are folded into the NUFFT's own load, interpolation
    and store passes.  bil ``[L, M]`` and cjl ``[L, N...]`` may be None (L = 1).
    ``coil_batch`` caps the coils per plan execution (0 = all); ``tol``
    defaults to 1e-6 in single precision (the reference's value) and 1e-13 in
    double precision (the reference asks for 1e-14; FP64 plans here go down
    to 1e-13, kernel width 14).
    """

    def __init__(self, coords: Sequence[torch.Tensor], smaps: torch.Tensor,
                 bil: Optional[torch.Tensor] = None, cjl: Optional[torch.Tensor] = None,
                 coil_batch: int = 0, tol: Optional[float] = None, **opt_kwargs):
        dtype = coords[0].dtype
        cdtype = torch.complex64 if dtype == torch.float32 else torch.complex128
        self.n = tuple(int(s) for s in smaps.shape[1:])
        self.C = int(smaps.shape[0])
        self.M = int(coords[0].shape[0])
        if len(coords) != len(self.n):
            raise RuntimeError("one coordinate vector per image axis is required")
        if (bil is None) != (cjl is None):
            raise RuntimeError("bil and cjl are given together")
        self.L = 1 if bil is None else int(bil.shape[0])
        for name, t, shape in (("smaps", smaps, None), ("bil", bil, (self.L, self.M)),
                               ("cjl", cjl, (self.L,) + self.n)):
            if t is None:
                continue
            hfft._require_gpu(t, name)
            if t.dtype != cdtype:
                raise RuntimeError(f"{name} must have dtype {cdtype}")
            if shape is not None and tuple(t.shape) != shape:
                raise RuntimeError(f"{name} must have shape {shape}, got {tuple(t.shape)}")
        self.smaps, self.bil, self.cjl = smaps, bil, cjl
        self.cdtype = cdtype
        nt = self.C if coil_batch <= 0 else min(int(coil_batch), self.C)
        if tol is None:
            tol = 1e-6 if dtype == torch.float32 else 1e-13
        mk = lambda sign, t: hfft.nufft_plan.make(  # noqa: E731
            hfft.nufft_opts(nmodes=self.n, ntransf=nt, tol=tol, sign=sign, **opt_kwargs), t, dtype)
        self.fwd = mk(hfft.nufft_sign.DEFAULT_TYPE_2, hfft.nufft_type.TYPE_2)
        self.adj = mk(hfft.nufft_sign.DEFAULT_TYPE_1, hfft.nufft_type.TYPE_1)
        self.fwd.setpts(coords)
        self.adj.setpts(coords)

    def forward(self, image: torch.Tensor, out: Optional[torch.Tensor] = None) -> torch.Tensor:
        hfft._require_gpu(image, "image")
        if image.dtype != self.cdtype or tuple(image.shape) != self.n:
            raise RuntimeError(f"image must be {self.cdtype} of shape {self.n}")
        if out is None:
            out = torch.empty((self.C, self.M), dtype=self.cdtype, device=image.device)
        hfft._require_gpu(out, "out")
        if out.dtype != self.cdtype or tuple(out.shape) != (self.C, self.M):
            raise RuntimeError(f"out must be {self.cdtype} of shape {(self.C, self.M)}")
        with torch.cuda.device(image.device):
            check(lib().hasty_nufft_set_stream(self.fwd._handle(), hfft._stream(image.device)))
            check(lib().hasty_sense_forward(self.fwd._handle(), self.C, self.L, image.data_ptr(),
                                            self.smaps.data_ptr(), hfft._ptr(self.bil),
                                            hfft._ptr(self.cjl), out.data_ptr()))
        return out

    def adjoint(self, kdata: torch.Tensor, combine: bool = False,
                out: Optional[torch.Tensor] = None) -> torch.Tensor:
        hfft._require_gpu(kdata, "kdata")
        if kdata.dtype != self.cdtype or tuple(kdata.shape) != (self.C, self.M):
            raise RuntimeError(f"kdata must be {self.cdtype} of shape {(self.C, self.M)}")
        shape = self.n if combine else (self.C,) + self.n
        if out is None:
            out = torch.empty(shape, dtype=self.cdtype, device=kdata.device)
        hfft._require_gpu(out, "out")
        if out.dtype != self.cdtype or tuple(out.shape) != shape:
            raise RuntimeError(f"out must be {self.cdtype} of shape {shape}")
        with torch.cuda.device(kdata.device):
            check(lib().hasty_nufft_set_stream(self.adj._handle(), hfft._stream(kdata.device)))
            check(lib().hasty_sense_adjoint(self.adj._handle(), self.C, self.L, kdata.data_ptr(),
                                            self.smaps.data_ptr(), hfft._ptr(self.bil),
                                            hfft._ptr(self.cjl), out.data_ptr(), int(bool(combine))))
        return out


def sense_forward_offresonance(coords, image, smaps, bil, cjl, coil_batch: int = 0,
                               tol: Optional[float] = None) -> torch.Tensor:
    """``forward.cppm:36-165``: kdata ``[C, M]`` = sum_l bil_l . A(cjl_l . S_c . image)."""
    return SenseOffresonance(coords, smaps, bil, cjl, coil_batch, tol).forward(image)


def sense_adjoint_offresonance(coords, kdata, smaps, bil, cjl, coil_batch: int = 0,
                               tol: Optional[float] = None) -> torch.Tensor:
    """``adjoint.cppm:33-171``: images ``[C, N...]`` =
    conj(S_c) . sum_l conj(cjl_l) . A^H(conj(bil_l) . kdata_c)."""
    return SenseOffresonance(coords, smaps, bil, cjl, coil_batch, tol).adjoint(kdata)


class SenseNufft(SenseOffresonance):
    """SENSE forward ``y_c = A(S_c x)`` and coil-combined adjoint
    ``sum_c conj(S_c) A^H y_c``: the L=1 case of :class:`SenseOffresonance`
    (products with S folded into the NUFFT passes, coil sum fused into the
    last FFT pass of the adjoint)."""

    def __init__(self, coords: Sequence[torch.Tensor], smaps: torch.Tensor, tol: float = 1e-5,
                 **opt_kwargs):
        super().__init__(coords, smaps, None, None, 0, tol, **opt_kwargs)

    def adjoint(self, y: torch.Tensor) -> torch.Tensor:  # type: ignore[override]
        return super().adjoint(y.contiguous(), combine=True)
