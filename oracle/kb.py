"""Kaiser-Bessel gridding NUFFT restated on the CPU -- test oracle only.

The reference delegates its NUFFT to cufinufft (``nufft.cppm:103-310``; the
library is fetched at an unpinned ``master`` tag, ``CMakeLists.txt:175-181``)
and no reference test pins its numbers, so this module restates the textbook
gridding algorithm (Jackson et al., IEEE TMI 1991; Beatty et al., IEEE TMI
2005) with the conventions of ``oracle.ndft``:

type 2 (``nufft.cppm:230-250``):
    deapodise the modes by 1/Phi, zero-pad into the fine grid at position
    ``n mod nf``, FFT with the transform's sign, interpolate with phi.
type 1 (``nufft.cppm:205-228``):
    spread with phi, FFT with the transform's sign, keep the centred modes,
    deapodise by 1/Phi.

phi(t)  = I0(beta sqrt(1 - (2t/w)^2))      (|t| <= w/2, t in fine-grid cells)
Phi(a)  = w sinh(sqrt(beta^2 - (a w/2)^2)) / sqrt(beta^2 - (a w/2)^2)
          (the window's continuous Fourier transform at angular frequency a
           = 2 pi n / nf; the sin form when the root is imaginary)
beta    = pi sqrt(w^2/sigma^2 (sigma - 1/2)^2 - 0.8)      (Beatty 2005, eq. 5)
width   = ceil(-log10(tol / 10)) at sigma = 2, else the smallest width whose
          aliasing estimate plus rounding floor is <= tol (``kernel_width``)

Every step is FP64.  The fine grid is ``nf_a = next 2^a 3^b 5^c >= sigma N_a``.
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np
import scipy.fft as sfft

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle_kb.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle gridding library missing: build it with `make -C {_HERE}`")
        lib = ctypes.CDLL(path)
        for name in ("kb_interp_f64", "kb_interp_f32", "kb_spread_f64", "kb_spread_f32"):
            getattr(lib, name).restype = None
        lib.kb_eval.restype = ctypes.c_double
        lib.kb_eval.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_double]
        _LIB = lib
    return _LIB


def kernel_beta(width: int, sigma: float = 2.0) -> float:
    """Beatty et al. 2005 shape parameter for width ``w`` at oversampling ``sigma``."""
    return math.pi * math.sqrt(width * width / (sigma * sigma) * (sigma - 0.5) ** 2 - 0.8)


def kernel_error_estimate(width: int, sigma: float = 2.0, dim: int = 3) -> float:
    """Aliasing estimate of the relative L2 gridding error.

    The window has compact support, so by Poisson summation the only error of
    a mode at angular frequency alpha (|alpha| <= pi/sigma) is the replicas
    ``Phi(alpha + 2 pi p)``, p != 0, relative to ``Phi(alpha)``.  RMS over 64
    midpoints of [0, pi/sigma] with |p| <= 400, times sqrt(dim).
    """
    beta = kernel_beta(width, sigma)
    al = np.pi / sigma * (np.arange(64) + 0.5) / 64
    p = np.concatenate([np.arange(1, 401), -np.arange(1, 401)]).astype(np.float64)
    with np.errstate(over="ignore", invalid="ignore"):
        rep = kb_hat(al[:, None] + 2 * np.pi * p[None, :], width, beta)
        h0 = kb_hat(al, width, beta)
    e2 = (rep ** 2).sum(1) / h0 ** 2
    return float(np.sqrt(dim * e2.mean()))


def kernel_rounding_gain(width: int, sigma: float = 2.0) -> float:
    """Per-axis growth of fine-grid rounding noise through the gridding steps:
    RMS deapodisation factor sqrt(mean Phi(alpha)^-2) over [0, pi/sigma] times
    the window's sqrt(int phi^2) (midpoint rule, 4096 points)."""
    beta = kernel_beta(width, sigma)
    al = np.pi / sigma * (np.arange(64) + 0.5) / 64
    inv = np.sqrt(np.mean(kb_hat(al, width, beta) ** -2.0))
    t = -0.5 * width + width * (np.arange(4096) + 0.5) / 4096
    phi2 = np.sum(kb_window(t, width, beta) ** 2) * width / 4096
    return float(inv * np.sqrt(phi2))


def kernel_width(tol: float, sigma: float = 2.0, dim: int = 3, precision: str = "f32") -> int:
    """Window width in fine-grid cells for a requested tolerance, for a plan
    computing in ``precision`` ("f32" for complex64 data, "f64").

    sigma = 2: FINUFFT's published rule ``ceil(-log10(tol / 10))``; any other
    sigma: the smallest width whose ``kernel_error_estimate`` plus rounding
    floor ``4 eps gain^dim`` (eps = 2^-24 or 2^-53) is <= tol.  17 means no
    width up to 16 reaches tol.
    """
    if sigma == 2.0:
        w = math.ceil(-math.log10(tol / 10.0))
        return 2 if w < 2 else (17 if w > 16 else int(w))
    eps = 2.0 ** -53 if precision == "f64" else 2.0 ** -24
    for w in range(2, 17):
        if kernel_error_estimate(w, sigma, dim) + 4 * eps * kernel_rounding_gain(w, sigma) ** dim <= tol:
            return w
    return 17


def next_smooth(n: int) -> int:
    """Smallest 2^a 3^b 5^c >= n."""
    m = max(int(n), 1)
    while True:
        k = m
        for p in (2, 3, 5):
            while k % p == 0:
                k //= p
        if k == 1:
            return m
        m += 1


def fine_size(n: int, sigma: float = 2.0, width: int = 6) -> int:
    return next_smooth(max(int(math.ceil(sigma * n)), 2 * width))


def kb_window(t, width: int, beta: float):
    t = np.asarray(t, np.float64)
    z = 1.0 - (2.0 * t / width) ** 2
    return np.where(z >= 0, np.i0(beta * np.sqrt(np.maximum(z, 0.0))), 0.0)


def kb_hat(alpha, width: int, beta: float):
    """Continuous Fourier transform of ``kb_window`` at angular frequency alpha."""
    alpha = np.asarray(alpha, np.float64)
    a = width / 2.0
    s = beta * beta - (a * alpha) ** 2
    r = np.sqrt(np.abs(s))
    safe = np.where(r == 0, 1.0, r)
    pos = 2 * a * np.sinh(r) / safe
    neg = 2 * a * np.sin(r) / safe
    out = np.where(s > 0, pos, neg)
    return np.where(r == 0, 2 * a, out)


def deapod_factors(n: int, nf: int, width: int, beta: float) -> np.ndarray:
    """1/Phi(2 pi m / nf) for the centred modes m of an axis of length n."""
    m = np.arange(-(n // 2), (n + 1) // 2, dtype=np.float64)
    return 1.0 / kb_hat(2 * np.pi * m / nf, width, beta)


def fold_coords(k, nf: int) -> np.ndarray:
    """Angular coordinate (radians, any range) -> fine-grid coordinate in [0, nf)."""
    k = np.asarray(k, np.float64)
    t = k / (2 * np.pi)
    t = t - np.floor(t)
    u = t * nf
    return np.where(u >= nf, u - nf, u)


class KBPlan:
    """FP64 (or FP32-grid) CPU restatement of a type-1 / type-2 NUFFT plan.

    Mirrors the life cycle of ``nufft_plan`` (``nufft.cppm:107-203``): the
    constructor fixes modes / sign / tolerance, ``setpts`` registers
    coordinates, ``type1`` / ``type2`` execute.
    """

    def __init__(self, shape, tol: float = 1e-5, sigma: float = 2.0, dtype=np.complex128,
                 workers: int | None = None, width_for: str = "f32"):
        """``width_for``: the precision whose width rule picks the window
        (the library plan this one restates), independent of ``dtype``."""
        self.shape = tuple(int(s) for s in shape)
        self.d = len(self.shape)
        self.w = kernel_width(tol, sigma, self.d, width_for)
        if self.w > 16:
            raise RuntimeError(f"tolerance {tol:g} at upsampling {sigma:g} needs a kernel wider than 16")
        self.beta = kernel_beta(self.w, sigma)
        self.nf = tuple(fine_size(n, sigma, self.w) for n in self.shape)
        self.dtype = np.dtype(dtype)
        self.workers = workers or len(os.sched_getaffinity(0))
        self.deapod = [deapod_factors(n, nf, self.w, self.beta) for n, nf in zip(self.shape, self.nf)]
        self.u = None

    def setpts(self, coords):
        if len({len(c) for c in coords}) != 1:
            raise RuntimeError("coordinate arrays have different lengths")
        self.M = len(coords[0])
        self.u = [np.ascontiguousarray(fold_coords(c, nf)) for c, nf in zip(coords, self.nf)]
        return self

    # -- helpers -----------------------------------------------------------
    def _deapod_nd(self):
        fac = 1.0
        for a in range(self.d):
            sh = [1] * self.d
            sh[a] = -1
            fac = fac * self.deapod[a].reshape(sh)
        return fac

    def _mode_index(self):
        return np.ix_(*[np.arange(-(n // 2), (n + 1) // 2) % nf for n, nf in zip(self.shape, self.nf)])

    def _uptrs(self):
        u = list(self.u) + [self.u[0]] * (3 - self.d)
        return [x.ctypes.data_as(ctypes.c_void_p) for x in u]

    def _real(self):
        return np.float64 if self.dtype == np.complex128 else np.float32

    # -- transforms ---------------------------------------------------------
    def type2(self, f, sign: int = -1) -> np.ndarray:
        f = np.asarray(f)
        squeeze = f.ndim == self.d
        if squeeze:
            f = f[None]
        T = f.shape[0]
        g = np.zeros((T,) + self.nf, self.dtype)
        g[(slice(None),) + self._mode_index()] = (f * self._deapod_nd()).astype(self.dtype)
        fft = sfft.fftn if sign < 0 else (lambda a, axes, workers: sfft.ifftn(a, axes=axes, workers=workers) * np.prod(self.nf))
        G = np.ascontiguousarray(fft(g, axes=tuple(range(1, self.d + 1)), workers=self.workers).astype(self.dtype))
        out = np.empty((T, self.M), self.dtype)
        fn = _lib().kb_interp_f64 if self.dtype == np.complex128 else _lib().kb_interp_f32
        nf = (ctypes.c_int64 * 3)(*(list(self.nf) + [1] * (3 - self.d)))
        u0, u1, u2 = self._uptrs()
        fn(ctypes.c_int(self.d), nf, ctypes.c_int(T), G.ctypes.data_as(ctypes.c_void_p),
           ctypes.c_int64(self.M), u0, u1, u2, ctypes.c_int(self.w), ctypes.c_double(self.beta),
           out.ctypes.data_as(ctypes.c_void_p))
        return out[0] if squeeze else out

    def type1(self, c, sign: int = +1) -> np.ndarray:
        c = np.ascontiguousarray(np.asarray(c, self.dtype))
        squeeze = c.ndim == 1
        if squeeze:
            c = c[None]
        T = c.shape[0]
        g = np.empty((T,) + self.nf, self.dtype)
        fn = _lib().kb_spread_f64 if self.dtype == np.complex128 else _lib().kb_spread_f32
        nf = (ctypes.c_int64 * 3)(*(list(self.nf) + [1] * (3 - self.d)))
        u0, u1, u2 = self._uptrs()
        fn(ctypes.c_int(self.d), nf, ctypes.c_int(T), c.ctypes.data_as(ctypes.c_void_p),
           ctypes.c_int64(self.M), u0, u1, u2, ctypes.c_int(self.w), ctypes.c_double(self.beta),
           g.ctypes.data_as(ctypes.c_void_p))
        axes = tuple(range(1, self.d + 1))
        if sign > 0:
            G = sfft.ifftn(g, axes=axes, workers=self.workers) * np.prod(self.nf)
        else:
            G = sfft.fftn(g, axes=axes, workers=self.workers)
        f = G[(slice(None),) + self._mode_index()] * self._deapod_nd()
        f = f.astype(self.dtype)
        return f[0] if squeeze else f
