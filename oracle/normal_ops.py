"""Multi-kernel / whitened SENSE normal operators and time-segmented SENSE --
test oracle only (FP64 numpy; imported by the tests and nothing else).

Each function restates one reference operator in the reference's own order
of operations, on top of ``toeplitz.toeplitz_multiplication`` (the batched
Toeplitz product with the per-element load / store rules of
``toeplitz_load_3D.cu:18-79``) and the direct NDFT:

* ``normal_t_t2``   -- ``NORMAL_T_T2_OP`` (``normal.cppm:133-155``): one
                       product, io NONE, in_m1 MULT x, out_m1 MULT_CONJ x.
* ``normal_idt_t1`` -- ``NORMAL_IDT_T1_OP`` (``normal.cppm:218-266``): per
                       kernel l, input S, io MULT_CONJ, in_m1 MULT x,
                       in_m2 MULT R_l, out_m2 MULT_CONJ R_l, accumulating
                       for l > 0; then the sum over coils.
* ``normal_idt_t2`` -- ``NORMAL_IDT_T2_OP`` (``normal.cppm:322-369``): as
                       above with input Psi, io NONE, out_m1 MULT_CONJ x.
* ``normal_idtw_t1``-- ``NORMAL_IDTW_T1_OP`` (``normal.cppm:441-491`` and its
                       script ``:531-534``): the stack with io NONE and
                       out_m1 NONE, then stack = W @ stack, times conj(S),
                       summed over coils.
* ``normal_idtw_t2``-- ``NORMAL_IDTW_T2_OP`` (``normal.cppm:619-668`` and its
                       script ``:705-708``): the T2 stack, then W @ stack.
* ``sense_forward_offres`` -- ``sense_forward_offresonance``
                       (``forward.cppm:116-128``): y_c = sum_l bil_l .
                       A(cjl_l . S_c . x), A the type-2 NDFT (sign -1).
* ``sense_adjoint_offres`` -- ``sense_adjoint_offresonance``
                       (``adjoint.cppm:123-135``): x_c = conj(S_c) . sum_l
                       conj(cjl_l) . A^H(conj(bil_l) . y_c), A^H the type-1
                       NDFT (sign +1); conj(S) applied once (the branch at
                       ``:116-117`` applies it twice).

Independent checks of these restatements (the tests): the Toeplitz identity
T_l z = (1/N^d) A^H diag(w_l) A z for a kernel built with weights w_l turns
every normal operator into a direct NDFT sum; the whitened T1 operator is
Hermitian for Hermitian W; the time-segmented forward equals the dense sum
y_j = sum_n S_n x_n e^{-i k_j.n} sum_l bil[l, j] cjl[l, n]; the adjoint
satisfies sum_c <A_c x, y_c> = <x, sum_c x_c> (the adjoint identity; the
coil sum of the adjoint's per-coil images is the adjoint of the forward).
"""
from __future__ import annotations

import numpy as np

from .ndft import ndft_type1, ndft_type2
from .toeplitz import MULT, MULT_CONJ, NONE, toeplitz_multiplication


def _c(a):
    return np.asarray(a, np.complex128)


def normal_t_t2(psi, x, kernel) -> np.ndarray:
    psi = _c(psi)
    return toeplitz_multiplication(psi, np.zeros_like(psi), kernel, m1=_c(x), io_type=NONE,
                                   in_m1=MULT, out_m1=MULT_CONJ)


def _stack(inp, x, kernels, kerneldiags, io, out_m1):
    inp = _c(inp)
    out = np.zeros_like(inp)
    for l in range(len(kernels)):
        out = toeplitz_multiplication(inp, out, kernels[l], m1=_c(x), m2=_c(kerneldiags[l]),
                                      io_type=io, in_m1=MULT, out_m1=out_m1, in_m2=MULT,
                                      out_m2=MULT_CONJ, accumulate=l > 0)
    return out


def _check(kernels, kerneldiags):
    if len(kernels) != len(kerneldiags):
        raise RuntimeError("Number of kernels and kernel diagonals must match.")


def normal_idt_t1(x, smaps, kernels, kerneldiags) -> np.ndarray:
    _check(kernels, kerneldiags)
    return _stack(smaps, x, kernels, kerneldiags, MULT_CONJ, NONE).sum(0)


def normal_idt_t2(psi, x, kernels, kerneldiags) -> np.ndarray:
    _check(kernels, kerneldiags)
    return _stack(psi, x, kernels, kerneldiags, NONE, MULT_CONJ)


def normal_idtw_t1(x, smaps, kernels, kerneldiags, weights) -> np.ndarray:
    _check(kernels, kerneldiags)
    s = _c(smaps)
    stack = _stack(s, x, kernels, kerneldiags, NONE, NONE)
    C = s.shape[0]
    stack = (_c(weights) @ stack.reshape(C, -1)).reshape(stack.shape)
    return (stack * np.conj(s)).sum(0)


def normal_idtw_t2(psi, x, kernels, kerneldiags, weights) -> np.ndarray:
    _check(kernels, kerneldiags)
    stack = _stack(psi, x, kernels, kerneldiags, NONE, MULT_CONJ)
    B = stack.shape[0]
    return (_c(weights) @ stack.reshape(B, -1)).reshape(stack.shape)


def sense_forward_offres(x, smaps, coords, bil, cjl) -> np.ndarray:
    s, x, bil, cjl = _c(smaps), _c(x), _c(bil), _c(cjl)
    out = np.zeros((s.shape[0], len(coords[0])), np.complex128)
    for l in range(bil.shape[0]):
        out += bil[l][None] * ndft_type2(s * (x * cjl[l])[None], coords, sign=-1)
    return out


def sense_adjoint_offres(y, smaps, coords, bil, cjl) -> np.ndarray:
    s, y, bil, cjl = _c(smaps), _c(y), _c(bil), _c(cjl)
    acc = np.zeros(s.shape, np.complex128)
    for l in range(bil.shape[0]):
        acc += np.conj(cjl[l])[None] * ndft_type1(y * np.conj(bil[l])[None], coords, s.shape[1:], sign=+1)
    return np.conj(s) * acc
