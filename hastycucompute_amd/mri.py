"""SENSE operators (host mirror of ``hasty::NORMAL_T_T1_OP`` and the L=1 SENSE
forward / adjoint, ``cpp/corelib/lib/mri/{normal,forward,adjoint}.cppm``)."""
from __future__ import annotations

import ctypes
from typing import Sequence

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


class SenseNufft:
    """SENSE forward ``y_c = A(S_c x)`` and adjoint ``sum_c conj(S_c) A^H y_c``
    with all coils batched through one type-2 and one type-1 plan.

    The L=1 case of ``sense_forward_offresonance`` (``forward.cppm:35-165``)
    and ``sense_adjoint_offresonance`` (``adjoint.cppm:32-171``), without the
    latter's duplicated conj(S) factor and with the coil combine.
    """

    def __init__(self, coords: Sequence[torch.Tensor], smaps: torch.Tensor, tol: float = 1e-5,
                 **opt_kwargs):
        self.smaps = smaps
        self.n = tuple(int(s) for s in smaps.shape[1:])
        C = int(smaps.shape[0])
        self.C = C
        self.M = int(coords[0].shape[0])
        self.fwd = hfft.nufft_plan.make(hfft.nufft_opts(nmodes=self.n, ntransf=C, tol=tol,
                                                        sign=hfft.nufft_sign.DEFAULT_TYPE_2,
                                                        **opt_kwargs), hfft.nufft_type.TYPE_2)
        self.adj = hfft.nufft_plan.make(hfft.nufft_opts(nmodes=self.n, ntransf=C, tol=tol,
                                                        sign=hfft.nufft_sign.DEFAULT_TYPE_1,
                                                        **opt_kwargs), hfft.nufft_type.TYPE_1)
        self.fwd.setpts(coords)
        self.adj.setpts(coords)

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        img = (self.smaps * x[None]).contiguous()
        y = torch.empty((self.C, self.M), dtype=torch.complex64, device=x.device)
        self.fwd.execute(img, y)
        return y

    def adjoint(self, y: torch.Tensor) -> torch.Tensor:
        img = torch.empty((self.C,) + self.n, dtype=torch.complex64, device=y.device)
        self.adj.execute(y.contiguous(), img)
        return (torch.conj(self.smaps) * img).sum(0)
