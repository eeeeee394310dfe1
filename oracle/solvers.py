"""SENSE normal operator, conjugate gradient and power iteration -- test oracle only.

* ``sense_normal``   -- ``NORMAL_T_T1_OP::operator()`` (``normal.cppm:65-88``):
                        sum_c conj(S_c) . T(S_c x) with T = ``toeplitz_multiply``.
* ``sense_normal_direct`` -- the same operator through the unnormalised NDFT,
                        (1/N^d) sum_c conj(S_c) A_1(A_2(S_c x)); the identity
                        between the two is the Toeplitz embedding theorem and is
                        what pins the kernel's indexing and scale.
* ``conjugate_gradient`` -- ``cg.cppm:122-184``: r = b - A x, p = r,
                        rz = Re<r,r>; per iteration alpha = rz / Re<p,Ap>,
                        x += alpha p, r -= alpha Ap, rz' = Re<r,r>,
                        p = r + (rz'/rz) p; stop when sqrt(rz) < tol.
                        Inner products and scalars in FP64.
* ``preconditioned_conjugate_gradient`` -- ``cg.cppm:16-119``: r = b - A x,
                        z = P r, p = z, rz = Re<r,z>; inner iterations (while
                        sqrt(rz) >= tol): alpha = rz / Re<p,Ap>, x += alpha p,
                        r -= alpha Ap, z = P r, rz' = Re<r,z>, p = z on a
                        restart else z + (rz'/rz) p; a restart ends the inner
                        loop; after_restart(x) after every outer iteration.
* ``max_eig``        -- ``alg.cppm:13-30``: y = A x, lambda = ||y||, x = y / lambda.
"""
from __future__ import annotations

import numpy as np

from .ndft import ndft_type1, ndft_type2
from .toeplitz import toeplitz_multiply


def sense_normal(x, smaps, kernel, dtype=np.complex128, workers=None) -> np.ndarray:
    x = np.asarray(x, dtype)
    s = np.asarray(smaps, dtype)
    t = toeplitz_multiply(s * x[None], kernel, dtype=dtype, workers=workers)
    return np.sum(np.conj(s) * t, axis=0)


def sense_normal_direct(x, smaps, coords) -> np.ndarray:
    x = np.asarray(x, np.complex128)
    s = np.asarray(smaps, np.complex128)
    shape = x.shape
    y = ndft_type2(s * x[None], coords, sign=-1)
    back = ndft_type1(y, coords, shape, sign=+1)
    return np.sum(np.conj(s) * back, axis=0) / float(np.prod(shape))


def sense_forward(x, smaps, coords) -> np.ndarray:
    """y_c = A(S_c x): the L=1 case of ``sense_forward_offresonance`` (``forward.cppm:35-165``)."""
    s = np.asarray(smaps, np.complex128)
    return ndft_type2(s * np.asarray(x, np.complex128)[None], coords, sign=-1)


def sense_adjoint(y, smaps, coords) -> np.ndarray:
    """sum_c conj(S_c) A^H y_c: the L=1 SENSE adjoint with coil combine
    (``adjoint.cppm:32-171`` without its duplicated conj(S) factor)."""
    s = np.asarray(smaps, np.complex128)
    back = ndft_type1(np.asarray(y, np.complex128), coords, s.shape[1:], sign=+1)
    return np.sum(np.conj(s) * back, axis=0)


def conjugate_gradient(A, b, x0, max_iter: int, tol: float = 0.0, history=None):
    x = np.array(x0, np.complex128, copy=True)
    b = np.asarray(b, np.complex128)
    r = b - A(x)
    p = r.copy()
    rz = float(np.vdot(r, r).real)
    resid = np.sqrt(rz)
    for _ in range(max_iter):
        if resid < tol:
            break
        Ap = A(p)
        pAp = float(np.vdot(p, Ap).real)
        alpha = rz / pAp
        x = x + alpha * p
        r = r - alpha * Ap
        rznew = float(np.vdot(r, r).real)
        p = r + (rznew / rz) * p
        rz = rznew
        resid = np.sqrt(rz)
        if history is not None:
            history.append(resid)
    return x


def preconditioned_conjugate_gradient(A, P, b, x0, max_inner_iter: int, max_outer_iter: int,
                                      tol: float = 1e-6, should_restart=None, after_restart=None):
    x = np.array(x0, np.complex128, copy=True)
    b = np.asarray(b, np.complex128)
    r = b - A(x)
    z = P(r)
    p = z.copy()
    rz = float(np.vdot(r, z).real)
    resid = np.sqrt(rz)
    for _ in range(max_outer_iter):
        for _ in range(max_inner_iter):
            if resid < tol:
                break
            Ap = A(p)
            alpha = rz / float(np.vdot(p, Ap).real)
            x = x + alpha * p
            r = r - alpha * Ap
            z = P(r)
            restart = bool(should_restart(x)) if should_restart is not None else False
            rznew = float(np.vdot(r, z).real)
            p = z.copy() if restart else z + (rznew / rz) * p
            rz = rznew
            resid = np.sqrt(rz)
            if restart:
                break
        if after_restart is not None:
            after_restart(x)
    return x


def max_eig(A, x0, max_iter: int = 10):
    x = np.asarray(x0, np.complex128)
    lam = 0.0
    for _ in range(max_iter):
        y = A(x)
        lam = float(np.linalg.norm(y.ravel()))
        x = y / lam
    return lam, x
