"""GPU parity: Toeplitz kernel, Toeplitz products and the SENSE normal
operator through the C-ABI vs the FP64 oracle."""
import itertools

import numpy as np
import pytest
import torch

from oracle import toeplitz as ot
from oracle.ndft import rel_l2
from oracle.solvers import sense_normal, sense_normal_direct
from tests.gpu_util import crandn, cuda, host, rand_coords, report

pytestmark = pytest.mark.gpu

TOL = 2e-5


def _kernel_gpu(coords, shape, weights=None):
    from hastycucompute_amd import fft as hf
    k = torch.empty(tuple(2 * n for n in shape), dtype=torch.complex64, device="cuda")
    w = None if weights is None else cuda(weights.reshape(1, -1))
    hf.toeplitz_kernel([cuda(c) for c in coords], k, w)
    return k


@pytest.mark.parametrize("shape,M", [((8, 10, 6), 2000), ((16, 16, 16), 6000), ((24, 20), 3000)])
def test_toeplitz_kernel_vs_ndft(shape, M):
    rng = np.random.default_rng(1)
    coords = rand_coords(rng, len(shape), M)
    k = _kernel_gpu(coords, shape)
    report(f"toeplitz_kernel {shape}", rel_l2(host(k), ot.toeplitz_kernel(coords, shape)), 2e-6)


def test_toeplitz_kernel_weighted():
    rng = np.random.default_rng(2)
    shape, M = (10, 12, 8), 3000
    coords = rand_coords(rng, 3, M)
    w = crandn(rng, M)
    k = _kernel_gpu(coords, shape, w)
    report("toeplitz_kernel weighted", rel_l2(host(k), ot.toeplitz_kernel(coords, shape, w)), 2e-6)


@pytest.mark.parametrize("shape,C,M", [((10, 12, 8), 3, 3000), ((24, 20), 2, 2500)])
def test_reference_sequence_kernel_transform_multiplication(shape, C, M):
    """The reference's call sequence (``toeplitz.cppm:510-712``): toeplitz_kernel,
    then transform_toeplitz_kernel, then toeplitz_multiplication with the SENSE
    modes of NORMAL_T_T1_OP (input S_c, mult1 = x, in/out MULT_CONJ, in_m1 MULT),
    then the coil sum, against the FP64 NDFT normal operator.  The transform
    leaves this library's kernel as it is, bit for bit."""
    from hastycucompute_amd import fft as hf
    rng = np.random.default_rng(17)
    coords = rand_coords(rng, len(shape), M)
    S = crandn(rng, (C,) + shape)
    x = crandn(rng, shape)
    k = _kernel_gpu(coords, shape)
    before = k.clone()
    hf.transform_toeplitz_kernel(k, False)
    assert torch.equal(k, before)
    out = torch.empty((C,) + shape, dtype=torch.complex64, device="cuda")
    T = hf.ToeplitzMultType
    hf.toeplitz_multiplication(cuda(S), out, k, None, cuda(x), None, T.MULT_CONJ, T.MULT, T.NONE, T.NONE,
                               T.NONE, hf.ToeplitzAccumulateType.NONE)
    got = host(out).astype(np.complex128).sum(0)
    report(f"toeplitz_kernel -> transform -> multiplication + sum {shape} C={C}",
           rel_l2(got, sense_normal_direct(x, S, coords)), TOL)


@pytest.mark.parametrize("shape", [(8, 10, 6), (16, 12, 20), (32, 24), (5, 6, 3)])
def test_toeplitz_multiply(shape):
    from hastycucompute_amd import fft as hf
    rng = np.random.default_rng(3)
    K = crandn(rng, tuple(2 * n for n in shape))
    x = crandn(rng, (3,) + shape)
    g = cuda(x)
    hf.toeplitz_multiply(g, cuda(K))
    report(f"toeplitz_multiply {shape}", rel_l2(host(g), ot.toeplitz_multiply(x, K)), TOL)


def test_toeplitz_multiplication_all_modes():
    from hastycucompute_amd import fft as hf
    rng = np.random.default_rng(4)
    shape = (6, 8, 10)
    K = crandn(rng, tuple(2 * n for n in shape))
    inp = crandn(rng, (2,) + shape)
    m1 = crandn(rng, shape)
    m2 = crandn(rng, shape)
    gK, gin, gm1, gm2 = cuda(K), cuda(inp), cuda(m1), cuda(m2)
    modes = [0, 1, 2]
    for io, i1, o1, i2, o2, acc in itertools.product(modes, modes, modes, [0, 2], [1], [0, 1]):
        prev = crandn(rng, inp.shape)
        out = cuda(prev)
        hf.toeplitz_multiplication(gin, out, gK, None, gm1, gm2, io, i1, o1, i2, o2, acc)
        ref = ot.toeplitz_multiplication(inp, prev, K, m1, m2, io, i1, o1, i2, o2, bool(acc))
        err = rel_l2(host(out), ref)
        assert err <= TOL, f"modes {(io, i1, o1, i2, o2, acc)}: {err:.2e}"
    print("[parity] toeplitz_multiplication: 108 mode combinations within 2e-5", flush=True)


def test_toeplitz_multiplication_no_diags_and_scratch():
    from hastycucompute_amd import fft as hf
    rng = np.random.default_rng(5)
    shape = (8, 8, 8)
    K = crandn(rng, (16, 16, 16))
    inp = crandn(rng, (3,) + shape)
    nbytes = hf.toeplitz_scratch_bytes(shape, 3)
    scratch = torch.empty(nbytes // 8, dtype=torch.complex64, device="cuda")
    out = torch.empty((3,) + shape, dtype=torch.complex64, device="cuda")
    hf.toeplitz_multiplication(cuda(inp), out, cuda(K), scratch, None, None, 0, 1, 1, 1, 1, 0)
    report("toeplitz_multiplication scratch", rel_l2(host(out), ot.toeplitz_multiply(inp, K)), TOL)
    with pytest.raises(RuntimeError):
        hf.toeplitz_multiplication(cuda(inp), out, cuda(K), scratch[:10], None, None, 0, 0, 0, 0, 0, 0)
    with pytest.raises(RuntimeError):
        hf.toeplitz_multiplication(cuda(inp), out, cuda(K), None, None, None, 3, 0, 0, 0, 0, 0)


@pytest.mark.parametrize("shape,C,M", [((12, 16, 10), 3, 4000), ((32, 27), 4, 5000)])
def test_normal_op_vs_ndft(shape, C, M):
    from hastycucompute_amd.mri import NORMAL_T_T1_OP
    rng = np.random.default_rng(6)
    coords = rand_coords(rng, len(shape), M)
    S = crandn(rng, (C,) + shape)
    x = crandn(rng, shape)
    k = _kernel_gpu(coords, shape)
    op = NORMAL_T_T1_OP(k, cuda(S))
    y = host(op(cuda(x)))
    report(f"NORMAL_T_T1_OP {shape} C={C} vs NDFT normal", rel_l2(y, sense_normal_direct(x, S, coords)), TOL)
    op2 = NORMAL_T_T1_OP(k, cuda(S), coil_batch=2)
    y2 = host(op2(cuda(x)))
    report("NORMAL_T_T1_OP coil_batch=2", rel_l2(y2, sense_normal(x, S, host(k))), TOL)


def test_cartesian_full_sampling_is_scaled_identity():
    """All N^d Cartesian points k = 2 pi m / N: A^H A = N^d I, so T = I."""
    from hastycucompute_amd import fft as hf
    shape = (8, 6, 10)
    grids = np.meshgrid(*[2 * np.pi * (np.arange(n) - n // 2) / n for n in shape], indexing="ij")
    coords = [g.ravel().astype(np.float32) for g in grids]
    k = _kernel_gpu(coords, shape)
    rng = np.random.default_rng(7)
    x = crandn(rng, (2,) + shape)
    g = cuda(x)
    hf.toeplitz_multiply(g, k)
    report("Cartesian T = I", rel_l2(host(g), x), TOL)


# (n, nz, coils, coil_batch, HASTY_TOEP_GROUPS): n x n planes (n = 64, 96,
# 128) take the fused passes; the cases cover one coil per block (partial
# sums), several coils per block with and without partial sums, an empty
# trailing group, and the chunked accumulation of a second coil batch
FUSED_CASES = [(128, 4, 3, 0, None), (128, 4, 3, 2, None), (128, 4, 5, 0, "2"), (128, 4, 4, 0, "1"),
               (128, 3, 7, 0, "3"), (128, 2, 2, 0, "5"), (96, 3, 3, 0, None), (96, 4, 5, 2, "2"),
               (96, 2, 4, 0, "1"), (64, 4, 3, 0, None), (64, 3, 5, 0, "2")]


@pytest.mark.parametrize("n,nz,C,chunk,groups", FUSED_CASES)
def test_normal_op_fused_last_pass(n, nz, C, chunk, groups, monkeypatch):
    """The fused passes (x and y pad FFTs in one kernel writing x
    class-major; y and x inverse FFTs, crops, maps and coil sum in another; the
    kernel read through the same permutation) vs the FP64 oracle, with a
    random kernel."""
    from hastycucompute_amd.mri import NORMAL_T_T1_OP
    if groups is not None:
        monkeypatch.setenv("HASTY_TOEP_GROUPS", groups)
    shape = (nz, n, n)
    rng = np.random.default_rng(11 + nz + C)
    S = crandn(rng, (C,) + shape)
    x = crandn(rng, shape)
    K = crandn(rng, tuple(2 * n for n in shape))
    op = NORMAL_T_T1_OP(cuda(K), cuda(S), coil_batch=chunk)
    y = host(op(cuda(x)))
    report(f"fused passes n={n} nz={nz} C={C} chunk={chunk} groups={groups}",
           rel_l2(y, sense_normal(x, S, K)), TOL)


@pytest.mark.parametrize("n", [128, 96])
def test_normal_idt_fused_last_pass(n):
    """Two kernels with spatial diagonals R_l (R_l applied in the fused first
    pass, conj(R_l) in the fused last pass, the second kernel accumulated) on
    n x n planes."""
    from hastycucompute_amd import mri
    from oracle import normal_ops as no
    shape = (3, n, n)
    rng = np.random.default_rng(12)
    C, L = 3, 2
    S = crandn(rng, (C,) + shape)
    x = crandn(rng, shape)
    K = crandn(rng, (L,) + tuple(2 * n for n in shape))
    R = np.exp(1j * rng.uniform(-np.pi, np.pi, (L,) + shape)).astype(np.complex64)
    op = mri.NORMAL_IDT_T1_OP(cuda(K), cuda(R), cuda(S))
    report(f"fused passes NORMAL_IDT_T1_OP n={n} L=2", rel_l2(host(op(cuda(x))), no.normal_idt_t1(x, S, K, R)), TOL)
