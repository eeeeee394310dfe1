"""Time-segmentation interpolators for off-resonance (field-map) correction.

Mirror of ``make_offresonance_interpolators`` (``cpp/corelib/lib/mri/off_resonance.cppm:22-130``,
whose TorchScript body is never compiled in the reference) and of
``tseg_interpolators`` (``hasty_python/simul_op.py:11-34``).  The factor the
time-segmented SENSE operators approximate (``forward.cppm:13-16``) is

    E[j, n] = exp(z_n t_j)  ~=  sum_l B[l, j] Ct[l, n],

for a complex rate map z (e.g. z = -R2* + i*omega) and sample times t_j, with
L segment times tl = linspace(min t, max t, L) and Ct[l, n] = exp(z_n tl_l).
B is the histogram-weighted least-squares fit over the values the map takes
(Sutton, Noll and Fessler 2003; Fessler et al. 2005):

    B[:, j] = argmin_b  sum_k h_k |exp(zk_k t_j) - sum_l b_l exp(zk_k tl_l)|^2
            = pinv(diag(sqrt h) Ch^T) diag(sqrt h) exp(zk t_j),   Ch[l, k] = exp(zk_k tl_l),

with zk the histogram bin centres and h the counts.  ``b0_only`` bins Im(z)
only (zk = i * centre, a pure frequency map); otherwise a 2-D histogram over
(Re z, Im z) is used.  ``autocorrelate`` replaces the histogram by its
autocorrelation, whose support is the set of value differences z_n - z_m
(centres spanning [-(max - min), max - min]); that is the distribution met by
the normal operator A^H A, whose exponent is (z_n - z_m) t.

Conventions fixed here where the reference's sources disagree (its
TorchScript uses exp(-i z t) for Ct but bins Im(z), and simul_op.py uses
opposite signs for Ch and Ct): the exponent is +z t throughout, as in the
forward operator's documentation, and a value is binned to the nearest bin
centre (the reference truncates, which moves the maximum into the second-last
bin).  The fit is computed in float64 on the field map's device and returned
as complex64 (complex128 with ``dtype=torch.complex128``).
"""
from __future__ import annotations

from typing import Tuple

import torch


def _bin_index(v: torch.Tensor, vmin: float, vmax: float, nbins: int) -> torch.Tensor:
    if vmax <= vmin:
        return torch.zeros_like(v, dtype=torch.long)
    return torch.round((v - vmin) / (vmax - vmin) * (nbins - 1)).long().clamp_(0, nbins - 1)


def _centres(vmin: float, vmax: float, nbins: int, device) -> torch.Tensor:
    return torch.linspace(vmin, vmax, nbins, dtype=torch.float64, device=device)


def histogram(fieldmap: torch.Tensor, nbins: int, b0_only: bool = False,
              autocorrelate: bool = False) -> Tuple[torch.Tensor, torch.Tensor]:
    """Bin centres zk (complex128 [K]) and counts h (float64 [K]) of the map's values."""
    if nbins < 1:
        raise ValueError("nbin must be >= 1")
    z = fieldmap.detach().reshape(-1).to(torch.complex128)
    dev = z.device
    if b0_only:
        v = z.imag
        vmin, vmax = float(v.min()), float(v.max())
        h = torch.bincount(_bin_index(v, vmin, vmax, nbins), minlength=nbins).to(torch.float64)
        c = _centres(vmin, vmax, nbins, dev)
        if autocorrelate:
            # counts of differences between bin centres: full correlation
            h = torch.nn.functional.conv1d(h.view(1, 1, -1), h.view(1, 1, -1), padding=nbins - 1).view(-1)
            step = (vmax - vmin) / (nbins - 1) if nbins > 1 else 0.0
            c = torch.arange(-(nbins - 1), nbins, dtype=torch.float64, device=dev) * step
        return torch.complex(torch.zeros_like(c), c), h
    if autocorrelate:
        raise ValueError("autocorrelation is defined for b0_only (frequency) maps")
    re, im = z.real, z.imag
    rmin, rmax, imin, imax = float(re.min()), float(re.max()), float(im.min()), float(im.max())
    ir = _bin_index(re, rmin, rmax, nbins)
    ii = _bin_index(im, imin, imax, nbins)
    # integer counts (an FP64 index_add_ of ones contends on a few bins)
    h = torch.bincount(ir * nbins + ii, minlength=nbins * nbins).to(torch.float64)
    cr = _centres(rmin, rmax, nbins, dev)
    ci = _centres(imin, imax, nbins, dev)
    zk = torch.complex(cr.view(-1, 1).expand(nbins, nbins), ci.view(1, -1).expand(nbins, nbins)).reshape(-1)
    return zk, h


def make_offresonance_interpolators(fieldmap: torch.Tensor, times: torch.Tensor, L: int, nbin: int,
                                    b0_only: bool = False, autocorrelate: bool = False,
                                    dtype: torch.dtype = torch.complex64
                                    ) -> Tuple[torch.Tensor, torch.Tensor]:
    """(B [L, M], Ct [L, N...]) with exp(z_n t_j) ~= sum_l B[l, j] Ct[l, n].

    ``fieldmap``: complex rate map [N...] (DIM 2 or 3); ``times``: real [M]."""
    if fieldmap.dim() not in (2, 3):
        raise ValueError("fieldmap must have rank 2 or 3")
    if L < 1:
        raise ValueError("L must be >= 1")
    if times.dim() != 1 or times.numel() < 1:
        raise ValueError("times must be a non-empty vector")
    dev = fieldmap.device
    t = times.detach().to(device=dev, dtype=torch.float64)
    tl = torch.linspace(float(t.min()), float(t.max()), L, dtype=torch.float64, device=dev)
    zk, h = histogram(fieldmap, nbin, b0_only, autocorrelate)
    keep = h > 0
    zk, w = zk[keep], h[keep].sqrt().to(torch.complex128)
    ch = torch.exp(zk.view(1, -1) * tl.view(-1, 1))                    # [L, K]
    # the fit of column j depends on t_j alone, and a trajectory repeats its
    # readout times (every spoke or interleave has the same ones): fit each
    # distinct time once and gather
    tu, inv = torch.unique(t, return_inverse=True)
    rhs = w.view(-1, 1) * torch.exp(zk.view(-1, 1) * tu.view(1, -1))    # [K, U]
    B = (torch.linalg.pinv(w.view(-1, 1) * ch.T) @ rhs)[:, inv]        # [L, M]
    z = fieldmap.detach().to(torch.complex128)
    Ct = torch.exp(z.unsqueeze(0) * tl.view((L,) + (1,) * z.dim()))
    return B.to(dtype).contiguous(), Ct.to(dtype).contiguous()
