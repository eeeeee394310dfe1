"""Time-segmentation interpolators -- test oracle only (numpy, FP64).

Restates the histogram-weighted least-squares fit of the reference's
``make_offresonance_interpolators`` (``off_resonance.cppm:81-124``) and
``tseg_interpolators`` (``simul_op.py:11-34``), with the conventions of the
build (exponent +z t, nearest-centre binning; see the package module):

    E[j, n] = exp(z_n t_j) ~= sum_l B[l, j] Ct[l, n],  Ct[l, n] = exp(z_n tl_l),
    tl = linspace(min t, max t, L),
    B[:, j] = argmin sum_k h_k |exp(zk t_j) - sum_l b_l exp(zk tl_l)|^2.

The solve uses numpy.linalg.lstsq per time column (the package uses a
pseudo-inverse), so the two agree only if both formulate the same weighted
problem.  Checks that pin it (tests): a map whose values lie on the bin
centres and take at most L distinct values is reproduced exactly; the fit
error falls as L grows; the autocorrelated histogram has the support of the
value differences and total mass (sum h)^2.
"""
from __future__ import annotations

import numpy as np


def _bins(v, nbins):
    vmin, vmax = float(v.min()), float(v.max())
    if vmax <= vmin:
        idx = np.zeros(v.shape, np.int64)
    else:
        idx = np.clip(np.rint((v - vmin) / (vmax - vmin) * (nbins - 1)).astype(np.int64), 0, nbins - 1)
    return idx, np.linspace(vmin, vmax, nbins)


def histogram(fieldmap, nbins, b0_only=False, autocorrelate=False):
    z = np.asarray(fieldmap, np.complex128).ravel()
    if b0_only:
        idx, c = _bins(z.imag, nbins)
        h = np.bincount(idx, minlength=nbins).astype(np.float64)
        if autocorrelate:
            h = np.correlate(h, h, mode="full")
            step = (c[-1] - c[0]) / (nbins - 1) if nbins > 1 else 0.0
            c = np.arange(-(nbins - 1), nbins) * step
        return 1j * c, h
    ir, cr = _bins(z.real, nbins)
    ii, ci = _bins(z.imag, nbins)
    h = np.bincount(ir * nbins + ii, minlength=nbins * nbins).astype(np.float64)
    zk = (cr[:, None] + 1j * ci[None, :]).ravel()
    return zk, h


def interpolators(fieldmap, times, L, nbins, b0_only=False, autocorrelate=False):
    t = np.asarray(times, np.float64)
    tl = np.linspace(t.min(), t.max(), L)
    zk, h = histogram(fieldmap, nbins, b0_only, autocorrelate)
    keep = h > 0
    zk, w = zk[keep], np.sqrt(h[keep])
    A = w[:, None] * np.exp(zk[:, None] * tl[None, :])          # [K, L]
    rhs = w[:, None] * np.exp(zk[:, None] * t[None, :])         # [K, M]
    B = np.linalg.lstsq(A, rhs, rcond=None)[0]                  # [L, M]
    z = np.asarray(fieldmap, np.complex128)
    Ct = np.exp(z[None] * tl.reshape((L,) + (1,) * z.ndim))
    return B, Ct
