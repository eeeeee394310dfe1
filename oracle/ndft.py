"""FP64 direct non-uniform DFT (the NUFFT's defining sums) -- test oracle only.

Conventions restated from the reference:

* type 1 ("BACKWARD", adjoint, default sign +1) and type 2 ("FORWARD",
  default sign -1): ``nufft.cppm:42-55``.
* type 1 maps nonuniform data ``[ntransf, M]`` to modes ``[ntransf, N...]``;
  type 2 maps modes to nonuniform data: ``nufft.cppm:205-250``.
* modes are centred, ``n_a in [-floor(N_a/2), ceil(N_a/2) - 1]`` in increasing
  order (the library default ``modeord = 0``; the reference never sets it,
  ``nufft.cppm:117``).
* ``coords[a]`` pairs with tensor axis ``a`` (axis 0 slowest):
  ``nufft.cppm:193-199`` passes ``coords[2]`` as the library's fastest (x)
  axis.

The 3-D sums are evaluated with the separable factorisation
``exp(s i k.n) = prod_a exp(s i k_a n_a)`` in chunks of points; this is exact
arithmetic re-association, not an approximation.
"""
from __future__ import annotations

import numpy as np


def centred_modes(n: int) -> np.ndarray:
    """Centred mode indices of an axis of length ``n`` (FINUFFT ``modeord=0``)."""
    return np.arange(-(n // 2), (n + 1) // 2, dtype=np.float64)


def _factors(coords, shape, sign, j0, j1):
    return [np.exp(sign * 1j * np.outer(np.asarray(coords[a][j0:j1], np.float64),
                                        centred_modes(shape[a])))
            for a in range(len(shape))]


def ndft_type2(f, coords, sign: int = -1, chunk: int = 4096) -> np.ndarray:
    """c[t, j] = sum_n f[t, n] exp(sign * i * k_j . n)  (``nufft.cppm:230-250``).

    ``f``: complex array ``[N...]`` or ``[T, N...]``; ``coords``: sequence of
    ``d`` real arrays ``[M]``.  Returns complex128 ``[M]`` or ``[T, M]``.
    """
    d = len(coords)
    f = np.asarray(f, np.complex128)
    squeeze = f.ndim == d
    if squeeze:
        f = f[None]
    shape = f.shape[1:]
    M = len(coords[0])
    out = np.empty((f.shape[0], M), np.complex128)
    for j0 in range(0, M, chunk):
        j1 = min(M, j0 + chunk)
        e = _factors(coords, shape, sign, j0, j1)
        if d == 1:
            out[:, j0:j1] = f @ e[0].T
        elif d == 2:
            t = np.einsum('tyx,jx->tjy', f, e[1])
            out[:, j0:j1] = np.einsum('tjy,jy->tj', t, e[0])
        else:
            t = np.einsum('tzyx,jx->tjzy', f, e[2])
            t = np.einsum('tjzy,jy->tjz', t, e[1])
            out[:, j0:j1] = np.einsum('tjz,jz->tj', t, e[0])
    return out[0] if squeeze else out


def ndft_type1(c, coords, shape, sign: int = +1, chunk: int = 4096) -> np.ndarray:
    """f[t, n] = sum_j c[t, j] exp(sign * i * k_j . n)  (``nufft.cppm:205-228``).

    ``c``: complex ``[M]`` or ``[T, M]``.  Returns complex128 ``[N...]`` or
    ``[T, N...]``.
    """
    d = len(coords)
    c = np.asarray(c, np.complex128)
    squeeze = c.ndim == 1
    if squeeze:
        c = c[None]
    shape = tuple(int(s) for s in shape)
    out = np.zeros((c.shape[0],) + shape, np.complex128)
    M = len(coords[0])
    for j0 in range(0, M, chunk):
        j1 = min(M, j0 + chunk)
        e = _factors(coords, shape, sign, j0, j1)
        cc = c[:, j0:j1]
        if d == 1:
            out += cc @ e[0]
        elif d == 2:
            t = np.einsum('tj,jy->tjy', cc, e[0])
            out += np.einsum('tjy,jx->tyx', t, e[1])
        else:
            t = np.einsum('tj,jz->tjz', cc, e[0])
            t = np.einsum('tjz,jy->tjzy', t, e[1])
            out += np.einsum('tjzy,jx->tzyx', t, e[2])
    return out[0] if squeeze else out


def rel_l2(a, b) -> float:
    """Relative L2 error ||a - b|| / ||b|| evaluated in FP64."""
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    nb = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (nb if nb > 0 else 1.0))
