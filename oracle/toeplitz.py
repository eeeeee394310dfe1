"""Toeplitz embedding of the NUFFT normal operator -- test oracle only.

Restates, in FP64 numpy:

* ``psf_direct``        -- the point-spread lags psf[m] = sum_j w_j e^{+i k_j.m},
                           m in [-(N-1), N-1]^d, by direct NDFT (the quantity
                           ``toeplitz_kernel`` computes with a type-1 NUFFT at
                           2N-1 modes, ``nufft.cppm:399-426``).
* ``toeplitz_kernel``   -- embed the lags at kernel index i = m + N - 1 of a
                           (2N)^d grid (index 2N-1 stays zero), fftn, scale by
                           1/N^d (``nufft.cppm:429-440``).
* ``toeplitz_multiply`` -- fftn with zero padding at the end to 2N, multiply by
                           the kernel, normalised ifftn, crop [N-1, 2N-1)
                           (``nufft.cppm:448-479``).  With the kernel above this
                           is (1/N^d) A^H A x for the unnormalised type-2 A.
* ``toeplitz_multiplication`` -- the batched, diagonal-weighted form of
                           ``toeplitz.cppm:611-712`` with the per-element load
                           and store rules of ``toeplitz_load_3D.cu:18-79``:
                             load : s   = in[b] (.)m1 (.)m2          (input types)
                             store: out = crop(T s) (.)m1 (.)m2 (.)in[b] (+out)
                           where (.)_t is a*b for MULT(1), a*conj(b) for
                           MULT_CONJ(2), skipped for NONE(0).  The convolution
                           uses the ``toeplitz_multiply`` scale (the reference's
                           VkFFT path differs by a constant factor; that factor
                           is not reproduced).
"""
from __future__ import annotations

import numpy as np
import scipy.fft as sfft

from .ndft import ndft_type1

NONE, MULT, MULT_CONJ = 0, 1, 2


def psf_direct(coords, shape, weights=None) -> np.ndarray:
    """Lags psf[m] for m in [-(N_a-1), N_a-1] per axis, complex128 ``[2N-1 ...]``."""
    M = len(coords[0])
    w = np.ones(M, np.complex128) if weights is None else np.asarray(weights, np.complex128)
    return ndft_type1(w, coords, tuple(2 * n - 1 for n in shape), sign=+1)


def toeplitz_kernel(coords, shape, weights=None) -> np.ndarray:
    """Frequency-domain Toeplitz kernel ``[2N ...]`` (``nufft.cppm:386-441``)."""
    shape = tuple(int(n) for n in shape)
    lags = psf_direct(coords, shape, weights)
    k = np.zeros(tuple(2 * n for n in shape), np.complex128)
    k[tuple(slice(0, 2 * n - 1) for n in shape)] = lags
    return sfft.fftn(k) / float(np.prod(shape))


def toeplitz_multiply(x, kernel, dtype=np.complex128, workers=None) -> np.ndarray:
    """``nufft.cppm:448-479`` on a batch ``[B, N...]`` (or a single ``[N...]``).

    ``dtype`` selects the arithmetic (complex128 for the oracle, complex64
    when timing the CPU baseline); ``workers`` is scipy.fft's thread count."""
    x = np.asarray(x, dtype)
    kernel = np.asarray(kernel, dtype)
    d = kernel.ndim
    squeeze = x.ndim == d
    if squeeze:
        x = x[None]
    shape = x.shape[1:]
    axes = tuple(range(1, d + 1))
    mid = sfft.fftn(x, s=tuple(2 * n for n in shape), axes=axes, workers=workers)
    mid *= kernel[None]
    mid = sfft.ifftn(mid, axes=axes, workers=workers)
    out = mid[(slice(None),) + tuple(slice(n - 1, 2 * n - 1) for n in shape)]
    return out[0] if squeeze else out


def _mul(a, b, t):
    if b is None or t == NONE:
        return a
    if t == MULT:
        return a * b
    if t == MULT_CONJ:
        return a * np.conj(b)
    raise ValueError(f"invalid ToeplitzMultType {t}")


def toeplitz_multiplication(inp, out, kernel, m1=None, m2=None, io_type=NONE, in_m1=NONE,
                            out_m1=NONE, in_m2=NONE, out_m2=NONE, accumulate=False) -> np.ndarray:
    """Batched Toeplitz product with diagonal operands (``toeplitz.cppm:611-712``)."""
    inp = np.asarray(inp, np.complex128)
    res = np.array(out, np.complex128, copy=True)
    for b in range(inp.shape[0]):
        s = _mul(_mul(inp[b], m1, in_m1), m2, in_m2)
        t = toeplitz_multiply(s, kernel)
        t = _mul(_mul(t, m1, out_m1), m2, out_m2)
        t = _mul(t, inp[b], io_type)
        res[b] = t + res[b] if accumulate else t
    return res
