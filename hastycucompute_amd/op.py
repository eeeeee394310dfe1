"""Solvers over square operators (host mirror of ``hasty::conjugate_gradient``
``cg.cppm:122-184`` and ``hasty::max_eig`` ``alg.cppm:13-30``).

Vector updates, inner products and the CG scalars run in the library's FP64
reduction kernels; the scalars stay on the device and convergence is a device
flag, so the iteration loop issues no host synchronisation.
"""
from __future__ import annotations

from typing import Callable, Tuple

import torch

from . import fft as hfft
from ._lib import CG_STATE_DOUBLES, check, lib


def conjugate_gradient(A: Callable[[torch.Tensor], torch.Tensor], *args, **kwargs):
    """The reference's two overloads (``cg.cppm:16-119`` and ``:122-184``):

    * ``conjugate_gradient(A, max_iter=0, tol=1e-6, check_every=0)`` -- plain CG;
    * ``conjugate_gradient(A, P, max_inner_iter, max_outer_iter, tol=1e-6)`` --
      preconditioned CG with restart callbacks (chosen when the second
      argument is callable; see ``preconditioned_conjugate_gradient``).
    """
    if args and callable(args[0]):
        return preconditioned_conjugate_gradient(A, *args, **kwargs)
    return _plain_conjugate_gradient(A, *args, **kwargs)


def _apply_into(A, v: torch.Tensor, out: torch.Tensor) -> None:
    apply = getattr(A, "apply", None)
    if apply is not None:
        apply(v, out)
    else:
        out.copy_(A(v))


def preconditioned_conjugate_gradient(A: Callable[[torch.Tensor], torch.Tensor],
                                      P: Callable[[torch.Tensor], torch.Tensor],
                                      max_inner_iter: int, max_outer_iter: int, tol: float = 1e-6):
    """``cg.cppm:16-119``.  Returns ``pcg(x, b, should_restart_callback=None,
    after_restart_callback=None) -> x`` (``x`` updated in place).

    r = b - A x, z = P r, p = z, rz = Re<r, z>.  Each outer iteration runs up
    to ``max_inner_iter`` steps while sqrt(rz) >= tol: alpha = rz / Re<p, Ap>,
    x += alpha p, r -= alpha Ap, z = P r, rz' = Re<r, z>, p = z if
    ``should_restart_callback(x)`` is true (which also ends the inner loop),
    else z + (rz'/rz) p; then ``after_restart_callback(x)``.  Inner products
    and scalars are FP64 on the device.  Without a restart callback the
    convergence test is the device flag (no host sync per step, later steps
    are no-ops); with one, the callback runs on the host every step, as in
    the reference.  ``P`` may offer ``apply(r, out)`` like ``A``.
    (The reference's ``requires`` clause names ``output_tensor_rank_t``, which
    its operators do not define, so it cannot be instantiated; this form has
    no such clause.)
    """

    def pcg(x: torch.Tensor, b: torch.Tensor, should_restart_callback=None,
            after_restart_callback=None) -> torch.Tensor:
        hfft._require_gpu(x, "x")
        hfft._require_gpu(b, "b")
        if x.dtype != torch.complex64 or b.shape != x.shape:
            raise RuntimeError("x and b must be complex64 tensors of the same shape")
        n = x.numel()
        s = hfft._stream(x.device)
        state = torch.zeros(CG_STATE_DOUBLES, dtype=torch.float64, device=x.device)
        r = b - A(x)
        z = torch.empty_like(x)
        p = torch.empty_like(x)
        Ap = torch.empty_like(x)
        _apply_into(P, r, z)
        check(lib().hasty_pcg_init(r.data_ptr(), z.data_ptr(), p.data_ptr(), n, float(tol),
                                   state.data_ptr(), s))
        for _ in range(max_outer_iter):
            for _ in range(max_inner_iter):
                if should_restart_callback is not None and bool(state[5].item()):
                    break
                _apply_into(A, p, Ap)
                check(lib().hasty_pcg_update(x.data_ptr(), r.data_ptr(), p.data_ptr(), Ap.data_ptr(), n,
                                             state.data_ptr(), s))
                _apply_into(P, r, z)
                restart = bool(should_restart_callback(x)) if should_restart_callback is not None else False
                check(lib().hasty_pcg_direction(r.data_ptr(), z.data_ptr(), p.data_ptr(), n, int(restart),
                                                state.data_ptr(), s))
                if restart:
                    break
            if after_restart_callback is not None:
                after_restart_callback(x)
        return x

    return pcg


def _plain_conjugate_gradient(A: Callable[[torch.Tensor], torch.Tensor], max_iter: int = 0,
                              tol: float = 1e-6, check_every: int = 0):
    """Returns ``cg(x, b) -> x`` (updates ``x`` in place), the shape of the
    reference's returned lambda.  ``A`` maps a c64 GPU tensor to one of the
    same shape; ``A.apply(p, out)`` is used when available to avoid an
    allocation per iteration.  ``check_every > 0`` reads the residual on the
    host every that many iterations (the reference reads it every iteration,
    ``cg.cppm:176``) and stops early; 0 leaves the stop to the device flag.
    """

    def cg(x: torch.Tensor, b: torch.Tensor, history=None) -> torch.Tensor:
        hfft._require_gpu(x, "x")
        hfft._require_gpu(b, "b")
        if x.dtype != torch.complex64 or b.shape != x.shape:
            raise RuntimeError("x and b must be complex64 tensors of the same shape")
        dev = x.device
        n = x.numel()
        s = hfft._stream(dev)
        state = torch.zeros(CG_STATE_DOUBLES, dtype=torch.float64, device=dev)
        r = b - A(x)
        p = torch.empty_like(x)
        Ap = torch.empty_like(x)
        check(lib().hasty_cg_init(r.data_ptr(), p.data_ptr(), n, float(tol), state.data_ptr(), s))
        apply = getattr(A, "apply", None)
        for it in range(max_iter):
            if apply is not None:
                apply(p, Ap)
            else:
                Ap.copy_(A(p))
            check(lib().hasty_cg_step(x.data_ptr(), r.data_ptr(), p.data_ptr(), Ap.data_ptr(), n,
                                      state.data_ptr(), s))
            if history is not None:
                history.append(state[0].sqrt())
            if check_every and (it + 1) % check_every == 0:
                if bool(state[5].item()):
                    break
        return x

    return cg


def cg_residual(state: torch.Tensor) -> float:
    return float(state[0].sqrt().item())


def max_eig(A: Callable[[torch.Tensor], torch.Tensor], x0: torch.Tensor,
            max_iter: int = 10) -> Tuple[float, torch.Tensor]:
    """``alg.cppm:13-30``: y = A x, lambda = ||y||, x = y / lambda (FP64 norm)."""
    s = hfft._stream(x0.device)
    nrm = torch.zeros(2, dtype=torch.float64, device=x0.device)
    x = x0
    for _ in range(max_iter):
        y = A(x)
        check(lib().hasty_norm(y.data_ptr(), y.numel(), nrm.data_ptr(), s))
        check(lib().hasty_scale_by_device(y.data_ptr(), y.numel(), nrm.data_ptr(), 1, s))
        x = y
    return float(nrm[0].item()), x
