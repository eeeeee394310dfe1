"""``torch.ops.hasty`` -- the path's operators registered with torch.library.

Registered ops (all on ROCm device tensors; complex64 data, float32
coordinates, or complex128 / float64 for FP64 plans):

* ``hasty::nufft_type2(Tensor image, Tensor[] coords, float tol) -> Tensor``
  image ``[T, N...]`` -> ``[T, M]``, sign -1 (``nufft.cppm:230-250``);
* ``hasty::nufft_type1(Tensor kdata, Tensor[] coords, int[] nmodes, float tol) -> Tensor``
  kdata ``[T, M]`` -> ``[T, N...]``, sign +1 (``nufft.cppm:205-228``);
* ``hasty::toeplitz_normal(Tensor x, Tensor kernel, Tensor smaps) -> Tensor``
  ``sum_c conj(S_c) T(S_c x)`` (``NORMAL_T_T1_OP``, ``normal.cppm:43-93``);
* ``hasty::sense_cg(Tensor b, Tensor kernel, Tensor smaps, int iters) -> Tensor``
  ``iters`` CG steps on that normal operator from x = 0 (``cg.cppm:122-184``),
  FP64 scalars on the device.

The NUFFT ops carry autograd formulas (image / k-space gradients through
the adjoint transform, trajectory gradients through
``hasty_nufft_coord_grad``) and fake kernels, so they trace under
``torch.compile`` / ``torch.export``.  Plans are cached per (type, modes,
transforms, tolerance, precision, device) and reuse their point sort while
the coordinate tensors are unchanged (pointer and version).
"""
from __future__ import annotations

from typing import Dict, List, Sequence, Tuple

import torch

from . import fft as hf
from . import op as hop
from .autograd import NufftOperator
from .mri import NORMAL_T_T1_OP

_plans: Dict[Tuple, NufftOperator] = {}
_MAX_PLANS = 8


def _operator(nmodes: Sequence[int], ntransf: int, tol: float, dtype: torch.dtype,
              device: torch.device) -> NufftOperator:
    key = (tuple(int(n) for n in nmodes), int(ntransf), float(tol), dtype, device.index)
    op = _plans.get(key)
    if op is None:
        if len(_plans) >= _MAX_PLANS:
            _plans.pop(next(iter(_plans)))
        with torch.cuda.device(device):
            op = NufftOperator(nmodes, ntransf=ntransf, tol=tol, dtype=dtype)
        _plans[key] = op
    return op


def _real_dtype(t: torch.Tensor) -> torch.dtype:
    return torch.float64 if t.dtype == torch.complex128 else torch.float32


# ---------------------------------------------------------------------------
# NUFFT type 2
# ---------------------------------------------------------------------------
@torch.library.custom_op("hasty::nufft_type2", mutates_args=())
def nufft_type2(image: torch.Tensor, coords: List[torch.Tensor], tol: float) -> torch.Tensor:
    if image.dim() != len(coords) + 1:
        raise RuntimeError("image must be [ntransf, N...] with one coordinate vector per axis")
    op = _operator(image.shape[1:], image.shape[0], tol, _real_dtype(image), image.device)
    op.set_coords(coords)
    return op.forward(image.contiguous())


@nufft_type2.register_fake
def _(image, coords, tol):
    return image.new_empty((image.shape[0], coords[0].shape[0]))


def _type2_setup(ctx, inputs, output):
    image, coords, tol = inputs
    ctx.save_for_backward(image, *coords)
    ctx.tol = tol


def _type2_backward(ctx, g):
    image, *coords = ctx.saved_tensors
    gi = gk = None
    if ctx.needs_input_grad[0]:
        gi = nufft_type1(g.contiguous(), list(coords), list(image.shape[1:]), ctx.tol)
    if ctx.needs_input_grad[1]:
        op = _operator(image.shape[1:], image.shape[0], ctx.tol, _real_dtype(image), image.device)
        op.set_coords(coords)
        gk = op.coord_grad(image.contiguous(), g.contiguous())
    return gi, gk, None


torch.library.register_autograd("hasty::nufft_type2", _type2_backward, setup_context=_type2_setup)


# ---------------------------------------------------------------------------
# NUFFT type 1
# ---------------------------------------------------------------------------
@torch.library.custom_op("hasty::nufft_type1", mutates_args=())
def nufft_type1(kdata: torch.Tensor, coords: List[torch.Tensor], nmodes: List[int], tol: float) -> torch.Tensor:
    if kdata.dim() != 2 or len(nmodes) != len(coords):
        raise RuntimeError("kdata must be [ntransf, M] with one coordinate vector per mode axis")
    op = _operator(nmodes, kdata.shape[0], tol, _real_dtype(kdata), kdata.device)
    op.set_coords(coords)
    return op.adjoint(kdata.contiguous())


@nufft_type1.register_fake
def _(kdata, coords, nmodes, tol):
    return kdata.new_empty((kdata.shape[0],) + tuple(nmodes))


def _type1_setup(ctx, inputs, output):
    kdata, coords, nmodes, tol = inputs
    ctx.save_for_backward(kdata, *coords)
    ctx.nmodes, ctx.tol = list(nmodes), tol


def _type1_backward(ctx, g):
    kdata, *coords = ctx.saved_tensors
    gc = gk = None
    if ctx.needs_input_grad[0]:
        gc = nufft_type2(g.contiguous(), list(coords), ctx.tol)
    if ctx.needs_input_grad[1]:
        op = _operator(ctx.nmodes, kdata.shape[0], ctx.tol, _real_dtype(kdata), kdata.device)
        op.set_coords(coords)
        gk = op.coord_grad(g.contiguous(), kdata.contiguous())
    return gc, gk, None, None


torch.library.register_autograd("hasty::nufft_type1", _type1_backward, setup_context=_type1_setup)


# ---------------------------------------------------------------------------
# SENSE normal operator and CG
# ---------------------------------------------------------------------------
@torch.library.custom_op("hasty::toeplitz_normal", mutates_args=())
def toeplitz_normal(x: torch.Tensor, kernel: torch.Tensor, smaps: torch.Tensor) -> torch.Tensor:
    return NORMAL_T_T1_OP(kernel.contiguous(), smaps.contiguous())(x.contiguous())


@toeplitz_normal.register_fake
def _(x, kernel, smaps):
    return torch.empty_like(x)


@torch.library.custom_op("hasty::sense_cg", mutates_args=())
def sense_cg(b: torch.Tensor, kernel: torch.Tensor, smaps: torch.Tensor, iters: int) -> torch.Tensor:
    A = NORMAL_T_T1_OP(kernel.contiguous(), smaps.contiguous())
    x = torch.zeros_like(b)
    hop.conjugate_gradient(A, max_iter=int(iters), tol=0.0)(x, b.contiguous())
    return x


@sense_cg.register_fake
def _(b, kernel, smaps, iters):
    return torch.empty_like(b)
