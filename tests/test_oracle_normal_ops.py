"""CPU: pin the oracle of the multi-kernel / whitened normal operators and of
the time-segmented SENSE forward / adjoint (``oracle/normal_ops.py``).

The restatements follow the reference's sequence of Toeplitz products with
diagonal load / store modes.  They are checked here against quantities that
do not go through those modes:

* the Toeplitz identity T_l z = (1/N^d) A^H diag(w_l) A z, with kernels built
  from weights w_l, turns every operator into a direct NDFT sum written
  from the operator's mathematical definition (complex R_l, S, x, W, w_l, so
  a dropped or misplaced conjugate, a missing diagonal or a transposed W
  fails);
* the whitened T1 operator is Hermitian for Hermitian W and real w_l;
* the time-segmented forward equals a dense brute-force sum over an explicit
  (M x N) exponential matrix, and the adjoint satisfies the adjoint identity;
* L = 1 with unit factors reduces to the plain SENSE pair.
"""
import numpy as np
import pytest

from oracle import normal_ops as no
from oracle import toeplitz as ot
from oracle.ndft import centred_modes, ndft_type1, ndft_type2, rel_l2
from oracle.solvers import sense_adjoint, sense_forward


def crandn(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def setup(shape, C=3, L=2, M=60, seed=0):
    rng = np.random.default_rng(seed)
    d = len(shape)
    coords = [rng.uniform(-np.pi, np.pi, M) for _ in range(d)]
    w = [rng.uniform(0.2, 1.5, M) * np.exp(1j * rng.uniform(-1, 1, M)) for _ in range(L)]
    kernels = np.stack([ot.toeplitz_kernel(coords, shape, w[l]) for l in range(L)])
    R = crandn(rng, (L,) + shape)
    S = crandn(rng, (C,) + shape)
    x = crandn(rng, shape)
    return rng, coords, w, kernels, R, S, x


def Tdirect(z, coords, w, shape):
    """(1/N^d) A^H diag(w) A z by NDFT, batched over the leading axis."""
    return ndft_type1(w[None] * ndft_type2(z, coords, sign=-1), coords, shape, sign=+1) / np.prod(shape)


SHAPES = [(6, 5), (4, 3, 5)]


@pytest.mark.parametrize("shape", SHAPES)
def test_idt_t1_against_ndft(shape):
    _, coords, w, K, R, S, x = setup(shape)
    got = no.normal_idt_t1(x, S, K, R)
    ref = sum(np.sum(np.conj(S) * np.conj(R[l])[None] * Tdirect(R[l][None] * S * x[None], coords, w[l], shape), 0)
              for l in range(len(K)))
    assert rel_l2(got, ref) < 1e-11


@pytest.mark.parametrize("shape", SHAPES)
def test_idt_t2_and_t_t2_against_ndft(shape):
    rng, coords, w, K, R, S, x = setup(shape)
    psi = crandn(rng, (4,) + shape)
    got = no.normal_idt_t2(psi, x, K, R)
    ref = sum(np.conj(x)[None] * np.conj(R[l])[None] * Tdirect(R[l][None] * x[None] * psi, coords, w[l], shape)
              for l in range(len(K)))
    assert rel_l2(got, ref) < 1e-11
    got1 = no.normal_t_t2(psi, x, K[0])
    ref1 = np.conj(x)[None] * Tdirect(x[None] * psi, coords, w[0], shape)
    assert rel_l2(got1, ref1) < 1e-11


@pytest.mark.parametrize("shape", SHAPES)
def test_idtw_t1_against_ndft_and_hermitian(shape):
    rng, coords, w, K, R, S, x = setup(shape)
    C = S.shape[0]
    Wm = crandn(rng, (C, C))
    got = no.normal_idtw_t1(x, S, K, R, Wm)
    stack = [sum(np.conj(R[l]) * Tdirect((R[l] * S[j] * x)[None], coords, w[l], shape)[0] for l in range(len(K)))
             for j in range(C)]
    ref = sum(np.conj(S[i]) * Wm[i, j] * stack[j] for i in range(C) for j in range(C))
    assert rel_l2(got, ref) < 1e-11
    # Hermitian W and real positive w_l give a Hermitian operator
    A = crandn(rng, (C, C))
    Wh = A @ A.conj().T + C * np.eye(C)
    wr = [np.abs(v) for v in w]
    Kr = np.stack([ot.toeplitz_kernel(coords, shape, wr[l]) for l in range(len(K))])
    y = crandn(rng, shape)
    lhs = np.vdot(y, no.normal_idtw_t1(x, S, Kr, R, Wh))
    rhs = np.vdot(no.normal_idtw_t1(y, S, Kr, R, Wh), x)
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


@pytest.mark.parametrize("shape", SHAPES)
def test_idtw_t2_against_ndft(shape):
    rng, coords, w, K, R, S, x = setup(shape)
    B = 3
    psi = crandn(rng, (B,) + shape)
    Wm = crandn(rng, (B, B))
    got = no.normal_idtw_t2(psi, x, K, R, Wm)
    stack = sum(np.conj(x)[None] * np.conj(R[l])[None] * Tdirect(R[l][None] * x[None] * psi, coords, w[l], shape)
                for l in range(len(K)))
    ref = np.einsum("ij,j...->i...", Wm, stack)
    assert rel_l2(got, ref) < 1e-11


def test_kernel_count_mismatch_raises():
    _, coords, w, K, R, S, x = setup((6, 5))
    with pytest.raises(RuntimeError, match="Number of kernels and kernel diagonals must match."):
        no.normal_idt_t1(x, S, K, R[:1])


def dense_forward(x, S, coords, bil, cjl, shape):
    """y[c, j] = sum_n S[c, n] x[n] e^{-i k_j.n} sum_l bil[l, j] cjl[l, n], by an explicit matrix."""
    grids = np.meshgrid(*[centred_modes(n) for n in shape], indexing="ij")
    phase = sum(np.outer(coords[a], grids[a].ravel()) for a in range(len(shape)))
    E = np.exp(-1j * phase) * (bil.T @ cjl.reshape(cjl.shape[0], -1))
    return (S.reshape(S.shape[0], -1) * x.ravel()[None]) @ E.T


@pytest.mark.parametrize("shape", SHAPES)
def test_offres_forward_dense_and_adjoint_identity(shape):
    rng, coords, _, _, _, S, x = setup(shape, C=3, M=70, seed=3)
    L = 3
    M = len(coords[0])
    bil = crandn(rng, (L, M))
    cjl = crandn(rng, (L,) + shape)
    y = no.sense_forward_offres(x, S, coords, bil, cjl)
    assert rel_l2(y, dense_forward(x, S, coords, bil, cjl, shape)) < 1e-12
    g = crandn(rng, y.shape)
    xa = no.sense_adjoint_offres(g, S, coords, bil, cjl)
    assert xa.shape == S.shape
    lhs = np.vdot(y, g)
    rhs = np.vdot(x, xa.sum(0))
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def test_offres_reduces_to_plain_sense():
    rng, coords, _, _, _, S, x = setup((6, 5), C=2, M=50, seed=4)
    shape = x.shape
    ones_b = np.ones((1, len(coords[0])), np.complex128)
    ones_c = np.ones((1,) + shape, np.complex128)
    y = no.sense_forward_offres(x, S, coords, ones_b, ones_c)
    assert rel_l2(y, sense_forward(x, S, coords)) < 1e-13
    g = crandn(rng, y.shape)
    assert rel_l2(no.sense_adjoint_offres(g, S, coords, ones_b, ones_c).sum(0), sense_adjoint(g, S, coords)) < 1e-13
