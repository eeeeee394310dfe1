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

``sense_cg_solve`` differentiates through a SENSE-CG reconstruction, either
exactly for the truncated solve (reverse mode through the checkpointed CG
iterates: one normal-operator apply and two kernel-derivative
interpolations per iteration) or, for a converged solve, by the implicit
function theorem (one adjoint CG solve on the same Toeplitz operator).
"""
from __future__ import annotations

from typing import List, Sequence

import torch

from . import fft as hf
from . import op as hop
from ._lib import CG_STATE_DOUBLES, check, lib
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


class _SenseCGAdjoint(torch.autograd.Function):
    """x_K = K steps of CG on N(k) x = b(k) from x_0 = 0, differentiated in
    reverse mode through the iterations (the exact gradient of the truncated
    solve).  The forward pass checkpoints every iterate it needs -- p_i,
    q_i = N p_i, r_{i+1} and the FP64 scalars rz_i, pq_i, alpha_i, rz_{i+1},
    beta_i -- and the backward pass runs the adjoint recurrence of
      q = N p, pq = Re<p,q>, alpha = rz/pq, x' = x + alpha p, r' = r - alpha q,
      rz' = ||r'||^2, beta = rz'/rz, p' = r' + beta p
    from the last step to the first: one normal-operator apply per step for
    the vector adjoint, and for the trajectory
      d/dk Re<g_q, N(k) p> = (1/N^d)[Re<A S p, dA(S g_q)> + Re<A S g_q, dA(S p)>]
    through the kernel-derivative interpolation; finally b(k) = (1/N^d)
    sum_c conj(S_c) A^H y_c contributes (1/N^d) Re<y, dA(S g_b)>."""

    @staticmethod
    def forward(ctx, y, smaps, op, iters, kernel_tol, *coords):
        op.set_coords(coords)
        nvox = float(torch.tensor(op.nmodes).prod())
        dev = y.device
        N = _normal_operator(op, smaps, coords, kernel_tol, nvox)
        b = ((smaps.conj() * op.adjoint(y)).sum(0) / nvox).contiguous()
        x = torch.zeros_like(b)
        r = b.clone()
        p = torch.empty_like(b)
        q = torch.empty_like(b)
        K = int(iters)
        P, Q, R = [], [], []
        scal = torch.zeros((K, 5), dtype=torch.float64, device=dev)   # rz_i, pq_i, alpha_i, rz_i+1, beta_i
        if b.dtype == torch.complex64:
            # the library's CG step: FP64 scalars on the device, FP32 vectors
            n = b.numel()
            s = hf._stream(dev)
            state = torch.zeros(CG_STATE_DOUBLES, dtype=torch.float64, device=dev)
            check(lib().hasty_cg_init(r.data_ptr(), p.data_ptr(), n, 0.0, state.data_ptr(), s))
            for i in range(K):
                P.append(p.clone())
                scal[i, 0] = state[0]
                N.apply(p, q)
                Q.append(q.clone())
                check(lib().hasty_cg_step(x.data_ptr(), r.data_ptr(), p.data_ptr(), q.data_ptr(), n,
                                          state.data_ptr(), s))
                scal[i, 1:5] = state[1:5]
                R.append(r.clone())
        else:
            # complex128: the same recurrence in FP64 throughout
            p.copy_(r)
            rz = _rdot64(r, r)
            for i in range(K):
                P.append(p.clone())
                N.apply(p, q)
                Q.append(q.clone())
                pq = _rdot64(p, q)
                alpha = rz / pq
                x += alpha * p
                r -= alpha * q
                rz1 = _rdot64(r, r)
                beta = rz1 / rz
                p.mul_(beta).add_(r)
                scal[i] = torch.stack([rz, pq, alpha, rz1, beta])
                R.append(r.clone())
                rz = rz1
        ctx.op, ctx.N, ctx.nvox = op, N, nvox
        ctx.P, ctx.Q, ctx.R, ctx.scal = P, Q, R, scal
        ctx.save_for_backward(y, smaps, b, *coords)
        return x

    @staticmethod
    def backward(ctx, gx):
        y, smaps, b, *coords = ctx.saved_tensors
        op, N, V = ctx.op, ctx.N, ctx.nvox
        op.set_coords(coords)
        want_k = any(ctx.needs_input_grad[5:])
        P, Q, R, scal = ctx.P, ctx.Q, ctx.R, ctx.scal
        K = len(P)
        f32 = lambda t: t.to(gx.real.dtype)     # noqa: E731  (FP64 device scalar -> the vectors' real type)
        gx = gx.contiguous()
        gr = torch.zeros_like(gx)
        gp = torch.zeros_like(gx)
        grz = torch.zeros((), dtype=torch.float64, device=gx.device)
        gk = [torch.zeros_like(c, dtype=op.dtype) for c in coords] if want_k else None
        Np = torch.empty_like(gx)
        for i in range(K - 1, -1, -1):
            rz, pq, alpha, rz1, beta = scal[i, 0], scal[i, 1], scal[i, 2], scal[i, 3], scal[i, 4]
            p_i, q_i, r1 = P[i], Q[i], R[i]
            # p' = r' + beta p
            gr = gr + gp
            gbeta = _rdot64(gp, p_i)
            gp_i = gp * f32(beta)
            # beta = rz'/rz
            grz1 = grz + gbeta / rz
            grz_i = -gbeta * beta / rz
            # rz' = ||r'||^2
            gr = gr + r1 * f32(2 * grz1)
            # r' = r - alpha q
            galpha = -_rdot64(gr, q_i)
            gq = gr * f32(-alpha)
            # x' = x + alpha p
            galpha = galpha + _rdot64(gx, p_i)
            gp_i = gp_i + gx * f32(alpha)
            # alpha = rz / pq
            grz_i = grz_i + galpha / pq
            gpq = -galpha * alpha / pq
            # pq = Re<p, q>
            gp_i = gp_i + q_i * f32(gpq)
            gq = gq + p_i * f32(gpq)
            # q = N p
            N.apply(gq.contiguous(), Np)
            gp_i = gp_i + Np
            if want_k:
                Sg = (smaps * gq.unsqueeze(0)).contiguous()
                Sp = (smaps * p_i.unsqueeze(0)).contiguous()
                t1 = op.coord_grad(Sg, op.forward(Sp))
                t2 = op.coord_grad(Sp, op.forward(Sg))
                for a in range(len(gk)):
                    gk[a] += (t1[a] + t2[a]) / V
            gp, grz = gp_i, grz_i
            # gr and gx carry over unchanged (r = r' + alpha q, x = x' - alpha p)
        # r_0 = b, p_0 = r_0, rz_0 = ||r_0||^2
        gb = gr + gp + b * f32(2 * grz)
        Sb = (smaps * gb.unsqueeze(0)).contiguous()
        gy = op.forward(Sb) / V if ctx.needs_input_grad[0] else None
        if want_k:
            t = op.coord_grad(Sb, y.contiguous())
            for a in range(len(gk)):
                gk[a] += t[a] / V
        else:
            gk = [None] * len(coords)
        return (gy, None, None, None, None, *gk)


def _rdot64(a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """Re <a, b> accumulated in FP64 on the device (0-d float64 tensor): the
    library's FP64 reduction for complex64, torch's for complex128."""
    aa, bb = a.contiguous(), b.contiguous()
    if aa.dtype == torch.complex128:
        return torch.vdot(aa.reshape(-1), bb.reshape(-1)).real
    out = torch.empty(2, dtype=torch.float64, device=a.device)
    check(lib().hasty_dot(aa.data_ptr(), bb.data_ptr(), aa.numel(), out.data_ptr(), hf._stream(a.device)))
    return out[0]


class _PairNormal:
    """N x = (1/N^d) sum_c conj(S_c) A^H A (S_c x) through the NUFFT pair of a
    complex128 operator (the Toeplitz passes are complex64)."""

    def __init__(self, op, smaps, nvox):
        self.op, self.smaps, self.nvox = op, smaps.contiguous(), nvox

    def apply(self, x, out):
        Sx = (self.smaps * x.unsqueeze(0)).contiguous()
        out.copy_((self.smaps.conj() * self.op.adjoint(self.op.forward(Sx))).sum(0) / self.nvox)
        return out


def _normal_operator(op, smaps, coords, kernel_tol, nvox):
    """The normal operator of the solve: the Toeplitz NORMAL_T_T1_OP (FP64 PSF)
    for complex64 operators, the NUFFT pair for complex128 ones."""
    if op.cdtype == torch.complex128:
        return _PairNormal(op, smaps, nvox)
    kernel = torch.empty(tuple(2 * n for n in op.nmodes), dtype=torch.complex64, device=smaps.device)
    hf.toeplitz_kernel(list(coords), kernel, tol=kernel_tol)
    return NORMAL_T_T1_OP(kernel, smaps.contiguous())


def sense_cg_solve(y: torch.Tensor, smaps: torch.Tensor, op: NufftOperator, *coords: torch.Tensor,
                   iters: int = 25, kernel_tol: float = 1e-9, method: str = "adjoint") -> torch.Tensor:
    """x_hat = ``iters`` CG steps on the Toeplitz normal equations of
    ||A(k) x - y||^2 from x = 0, differentiable in y and k.

    ``method="adjoint"`` (default) differentiates the truncated solve exactly,
    in reverse mode through the checkpointed iterates; ``method="implicit"``
    returns the gradient of the converged solution (one adjoint CG solve,
    valid once the forward solve has converged).  A complex128 operator
    (``NufftOperator(dtype=torch.float64)``) runs the adjoint method in FP64
    throughout, on the normal operator of its NUFFT pair."""
    if method == "implicit" and op.cdtype != torch.complex64:
        raise RuntimeError("method 'implicit' needs a complex64 operator")
    if method == "adjoint":
        return _SenseCGAdjoint.apply(y, smaps, op, iters, kernel_tol, *coords)
    if method == "implicit":
        return _SenseCG.apply(y, smaps, op, iters, kernel_tol, *coords)
    raise RuntimeError("method must be 'adjoint' or 'implicit'")
