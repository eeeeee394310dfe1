"""Differentiable NUFFT and SENSE operators (torch.autograd) on the gfx950 path.

The reference's operators are not differentiable; this layer adds
gradients with respect to the image / k-space data and to the trajectory:

* ``nufft2(image, op, *coords)``: y = A(k) x (type 2, sign -1).
  Backward: dL/dx = A^H g (the type-1 adjoint NUFFT) and
  dL/dk_{j,a} = Re sum_t conj(g_t,j) d/dk_{j,a} y_t,j, computed on the GPU by
  interpolation with the derivative of the Kaiser-Bessel window
  (``hasty_nufft_coord_grad``).
* ``nufft1(kdata, op, *coords)``: x = A^H(k) c (type 1, sign +1).
  Backward: dL/dc = A g and dL/dk = Re sum_t conj(c_t) d/dk A(g_t).

PyTorch's convention for complex tensors is used: the gradient of a real
loss with respect to a complex tensor z is the conjugate Wirtinger
derivative scaled by 2, so for a linear map y = A x the input gradient is
A^H g, and for a real parameter k it is Re(conj(g) dy/dk).

``sense_cg_solve`` differentiates through a SENSE-CG reconstruction by the
implicit-function theorem: one adjoint CG solve on the same Toeplitz normal
operator instead of backpropagating through the iterations.  Only the
converged iterate is checkpointed for the backward pass (the operator and
its PSF are kept), so memory does not grow with the iteration count.
"""
from __future__ import annotations

from typing import List, Sequence

import torch

from . import fft as hf
from . import op as hop
from .mri import NORMAL_T_T1_OP


class NufftOperator:
    """A pair of plans (type 2 sign -1 and type 1 sign +1) sharing one
    trajectory, for ``ntransf`` transforms of shape ``nmodes``.  Coordinates
    are (re)registered only when they change (pointer and version)."""

    def __init__(self, nmodes: Sequence[int], ntransf: int = 1, tol: float = 1e-5,
                 dtype: torch.dtype = torch.float32, **opt_kwargs):
        self.nmodes = tuple(int(n) for n in nmodes)
        self.ntransf = int(ntransf)
        self.dtype = dtype
        self.cdtype = torch.complex64 if dtype == torch.float32 else torch.complex128
        mk = lambda sign, nt: hf.nufft_plan.make(  # noqa: E731
            hf.nufft_opts(nmodes=self.nmodes, ntransf=self.ntransf, tol=tol, sign=sign, **opt_kwargs), nt, dtype)
        self.fwd = mk(hf.nufft_sign.DEFAULT_TYPE_2, hf.nufft_type.TYPE_2)
        self.adj = mk(hf.nufft_sign.DEFAULT_TYPE_1, hf.nufft_type.TYPE_1)
        self._key = None
        self.M = None

    def set_coords(self, coords: Sequence[torch.Tensor]) -> None:
        coords = [c.detach().contiguous() for c in coords]
        key = tuple((c.data_ptr(), c._version, int(c.shape[0])) for c in coords)
        if key == self._key:
            return
        self.fwd.setpts(coords)
        self.adj.setpts(coords)
        self._key = key
        self._coords = coords
        self.M = int(coords[0].shape[0])

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        y = torch.empty((self.ntransf, self.M), dtype=self.cdtype, device=x.device)
        self.fwd.execute(x.contiguous(), y)
        return y

    def adjoint(self, y: torch.Tensor) -> torch.Tensor:
        x = torch.empty((self.ntransf,) + self.nmodes, dtype=self.cdtype, device=y.device)
        self.adj.execute(y.contiguous(), x)
        return x

    def coord_grad(self, h: torch.Tensor, w: torch.Tensor) -> List[torch.Tensor]:
        return self.fwd.coord_grad(h.contiguous(), w.contiguous())


class _Nufft2(torch.autograd.Function):
    @staticmethod
    def forward(ctx, image, op, *coords):
        op.set_coords(coords)
        ctx.op = op
        ctx.save_for_backward(image, *coords)
        return op.forward(image)

    @staticmethod
    def backward(ctx, g):
        image, *coords = ctx.saved_tensors
        op = ctx.op
        op.set_coords(coords)
        g = g.contiguous()
        gx = op.adjoint(g) if ctx.needs_input_grad[0] else None
        gk = [None] * len(coords)
        if any(ctx.needs_input_grad[2:]):
            gk = op.coord_grad(image, g)
        return (gx, None, *gk)


class _Nufft1(torch.autograd.Function):
    @staticmethod
    def forward(ctx, kdata, op, *coords):
        op.set_coords(coords)
        ctx.op = op
        ctx.save_for_backward(kdata, *coords)
        return op.adjoint(kdata)

    @staticmethod
    def backward(ctx, g):
        kdata, *coords = ctx.saved_tensors
        op = ctx.op
        op.set_coords(coords)
        g = g.contiguous()
        gc = op.forward(g) if ctx.needs_input_grad[0] else None
        gk = [None] * len(coords)
        if any(ctx.needs_input_grad[2:]):
            gk = op.coord_grad(g, kdata)
        return (gc, None, *gk)


def nufft2(image: torch.Tensor, op: NufftOperator, *coords: torch.Tensor) -> torch.Tensor:
    """y[t, j] = sum_n image[t, n] exp(-i k_j . n), differentiable in image and k."""
    return _Nufft2.apply(image, op, *coords)


def nufft1(kdata: torch.Tensor, op: NufftOperator, *coords: torch.Tensor) -> torch.Tensor:
    """x[t, n] = sum_j kdata[t, j] exp(+i k_j . n), differentiable in kdata and k."""
    return _Nufft1.apply(kdata, op, *coords)


def sense_forward(x: torch.Tensor, smaps: torch.Tensor, op: NufftOperator,
                  *coords: torch.Tensor) -> torch.Tensor:
    """y_c = A(k) (S_c x) for all coils (ntransf of ``op`` = coils)."""
    return nufft2(smaps * x.unsqueeze(0), op, *coords)


def sense_adjoint(y: torch.Tensor, smaps: torch.Tensor, op: NufftOperator,
                  *coords: torch.Tensor) -> torch.Tensor:
    return (smaps.conj() * nufft1(y, op, *coords)).sum(0)


# ---------------------------------------------------------------------------
# Gradient through a SENSE-CG solve (implicit adjoint)
# ---------------------------------------------------------------------------
class _SenseCG(torch.autograd.Function):
    """x_hat = N(k)^{-1} b(k) with N = (1/N^d) A^H A, b = (1/N^d) A^H y, solved
    by CG on the Toeplitz normal operator.  Backward: lambda = N^{-1} g (one
    CG solve, N is Hermitian), then with r = y - A x_hat,
      dL/dk = (1/N^d) [Re(conj(r) dA(lambda)/dk) - Re(conj(A lambda) dA(x_hat)/dk)]
    summed over coils (dA applied to S_c lambda and S_c x_hat)."""

    @staticmethod
    def forward(ctx, y, smaps, op, iters, kernel_tol, *coords):
        op.set_coords(coords)
        nvox = float(torch.tensor(op.nmodes).prod())
        kernel = torch.empty(tuple(2 * n for n in op.nmodes), dtype=torch.complex64, device=y.device)
        hf.toeplitz_kernel(list(coords), kernel, tol=kernel_tol)
        N = NORMAL_T_T1_OP(kernel, smaps.contiguous())
        b = (smaps.conj() * op.adjoint(y)).sum(0) / nvox
        x = torch.zeros_like(b)
        cg = hop.conjugate_gradient(N, max_iter=int(iters), tol=0.0)
        cg(x, b.contiguous(), history=None)
        ctx.op, ctx.N, ctx.iters, ctx.nvox = op, N, int(iters), nvox
        ctx.save_for_backward(y, smaps, x, *coords)    # checkpoint: the solution
        return x

    @staticmethod
    def backward(ctx, g):
        y, smaps, x, *coords = ctx.saved_tensors
        op, N = ctx.op, ctx.N
        op.set_coords(coords)
        lam = torch.zeros_like(x)
        hop.conjugate_gradient(N, max_iter=ctx.iters, tol=0.0)(lam, g.contiguous())
        gy = None
        if ctx.needs_input_grad[0]:
            # x_hat is linear in y: dL/dy = (1/N^d) A (S lambda)
            gy = op.forward(smaps * lam.unsqueeze(0)) / ctx.nvox
        gk = [None] * len(coords)
        if any(ctx.needs_input_grad[5:]):
            Sl = (smaps * lam.unsqueeze(0)).contiguous()
            Sx = (smaps * x.unsqueeze(0)).contiguous()
            r = (y - op.forward(Sx)).contiguous()
            Al = op.forward(Sl)
            t1 = op.coord_grad(Sl, r)
            t2 = op.coord_grad(Sx, Al)
            gk = [(a - b) / ctx.nvox for a, b in zip(t1, t2)]
        return (gy, None, None, None, None, *gk)


def sense_cg_solve(y: torch.Tensor, smaps: torch.Tensor, op: NufftOperator, *coords: torch.Tensor,
                   iters: int = 25, kernel_tol: float = 1e-9) -> torch.Tensor:
    """x_hat = argmin ||A(k) x - y||^2 by ``iters`` CG steps on the Toeplitz
    normal operator; differentiable in y and k (implicit adjoint)."""
    return _SenseCG.apply(y, smaps, op, iters, kernel_tol, *coords)
