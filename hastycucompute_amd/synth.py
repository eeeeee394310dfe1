"""Deterministic synthetic inputs for the reconstruction workloads.

* image       -- complex Gaussian, Re, Im ~ N(0, 1/2)
* coil maps   -- C smooth Gaussian bumps (centres uniform in the field of
                 view, width ~ N/4, random complex phase), normalised by the
                 root-sum-of-squares over coils
* kooshball   -- 3-D radial spokes through the k-space centre, samples
                 r_t = -pi + 2 pi t / T, directions from the 2-D golden means
                 (Chan et al., MRM 2009): cos(theta) = frac(s phi1),
                 phi = 2 pi frac(s phi2), phi1 = 0.4656, phi2 = 0.6823
* spirals     -- 2-D variable-density Archimedean spiral interleaves

Coordinates are float32 radians in [-pi, pi), ``coords[a]`` pairs with
image axis ``a`` (axis 0 slowest).
"""
from __future__ import annotations

import math

import numpy as np

PHI1 = 0.4656
PHI2 = 0.6823


def image(shape, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    s = math.sqrt(0.5)
    return (rng.standard_normal(shape, np.float32) * s
            + 1j * rng.standard_normal(shape, np.float32) * s).astype(np.complex64)


def coil_maps(shape, ncoils: int, seed: int = 1) -> np.ndarray:
    rng = np.random.default_rng(seed)
    d = len(shape)
    grids = np.meshgrid(*[np.arange(n, dtype=np.float32) for n in shape], indexing="ij")
    maps = np.empty((ncoils,) + tuple(shape), np.complex64)
    for c in range(ncoils):
        centre = [rng.uniform(0, n) for n in shape]
        width = [n / 4.0 for n in shape]
        phase = rng.uniform(0, 2 * np.pi)
        r2 = np.zeros(shape, np.float32)
        for a in range(d):
            r2 += ((grids[a] - centre[a]) / width[a]) ** 2
        maps[c] = (np.exp(-0.5 * r2) * np.exp(1j * phase)).astype(np.complex64)
    rss = np.sqrt(np.sum(np.abs(maps) ** 2, axis=0))
    maps /= np.maximum(rss, 1e-12)[None]
    return maps


def kooshball(nspokes: int, nsamples: int, offset: int = 0):
    """3 float32 arrays [nspokes * nsamples] (axis 0, 1, 2)."""
    s = np.arange(offset, offset + nspokes, dtype=np.float64)
    cos_t = np.mod(s * PHI1, 1.0)
    phi = 2 * np.pi * np.mod(s * PHI2, 1.0)
    sin_t = np.sqrt(1.0 - cos_t ** 2)
    r = -np.pi + 2 * np.pi * np.arange(nsamples, dtype=np.float64) / nsamples
    kz = np.outer(cos_t, r)
    ky = np.outer(sin_t * np.sin(phi), r)
    kx = np.outer(sin_t * np.cos(phi), r)
    return [np.ascontiguousarray(a.ravel().astype(np.float32)) for a in (kz, ky, kx)]


def rotated_kooshball(nspokes: int, nsamples: int, frame: int):
    """Frame ``frame`` of a dynamic acquisition: the kooshball rotated about
    axis 0 by the golden angle per frame."""
    kz, ky, kx = kooshball(nspokes, nsamples, offset=frame * 7919)
    ang = frame * math.pi * (3 - math.sqrt(5))
    c, s = math.cos(ang), math.sin(ang)
    ky2 = c * ky.astype(np.float64) - s * kx.astype(np.float64)
    kx2 = s * ky.astype(np.float64) + c * kx.astype(np.float64)
    wrap = lambda a: (np.mod(a + np.pi, 2 * np.pi) - np.pi).astype(np.float32)
    return [kz, wrap(ky2), wrap(kx2)]


def spiral2d(ninterleaves: int, nsamples: int, density_power: float = 1.5):
    """Variable-density spiral interleaves, float32 [ninterleaves * nsamples] x 2."""
    t = np.linspace(0.0, 1.0, nsamples, endpoint=False)
    r = np.pi * t ** density_power
    turns = 16.0 / ninterleaves * 8
    th = 2 * np.pi * turns * t
    out_y, out_x = [], []
    for i in range(ninterleaves):
        a = th + 2 * np.pi * i / ninterleaves
        out_y.append(r * np.sin(a))
        out_x.append(r * np.cos(a))
    y = np.clip(np.concatenate(out_y), -np.pi, np.pi - 1e-6)
    x = np.clip(np.concatenate(out_x), -np.pi, np.pi - 1e-6)
    return [y.astype(np.float32), x.astype(np.float32)]


def ellipse_phantom(shape, nellipses: int = 10, seed: int = 4, texture: float = 0.1) -> np.ndarray:
    """2-D complex test image: a sum of random ellipses plus a smooth
    Gaussian texture (complex64)."""
    rng = np.random.default_rng(seed)
    ny, nx = shape
    yy, xx = np.meshgrid(np.linspace(-1, 1, ny), np.linspace(-1, 1, nx), indexing="ij")
    img = np.zeros(shape, np.complex128)
    for _ in range(nellipses):
        cy, cx = rng.uniform(-0.5, 0.5, 2)
        ay, ax = rng.uniform(0.1, 0.45, 2)
        th = rng.uniform(0, np.pi)
        c, s = np.cos(th), np.sin(th)
        u = (c * (xx - cx) + s * (yy - cy)) / ax
        v = (-s * (xx - cx) + c * (yy - cy)) / ay
        img += rng.uniform(0.2, 1.0) * np.exp(1j * rng.uniform(0, 2 * np.pi)) * (u * u + v * v <= 1.0)
    noise = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    kern = np.exp(-0.5 * ((yy / 0.05) ** 2 + (xx / 0.05) ** 2))
    smooth = np.fft.ifft2(np.fft.fft2(noise) * np.fft.fft2(np.fft.ifftshift(kern)))
    img += texture * smooth / np.abs(smooth).max()
    return img.astype(np.complex64)


def stack_of_spirals(npartitions: int, ninterleaves: int, nsamples: int, density_power: float = 1.5):
    """3-D stack of spirals: ``npartitions`` k_z planes in [-pi, pi), each
    with ``ninterleaves`` variable-density spiral interleaves in (k_y, k_x);
    float32 arrays (axis 0, 1, 2)."""
    ky, kx = spiral2d(ninterleaves, nsamples, density_power)
    kz_planes = -np.pi + 2 * np.pi * np.arange(npartitions) / npartitions
    kz = np.repeat(kz_planes, ky.size).astype(np.float32)
    return [kz, np.tile(ky, npartitions), np.tile(kx, npartitions)]
