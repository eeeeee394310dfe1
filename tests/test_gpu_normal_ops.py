"""GPU parity: the multi-kernel / whitened normal operators
(NORMAL_T_T2_OP, NORMAL_IDT_T1_OP, NORMAL_IDT_T2_OP, NORMAL_IDTW_T1_OP,
NORMAL_IDTW_T2_OP) and the time-segmented SENSE forward / adjoint through
the C-ABI, against the FP64 oracle on the same seeded inputs."""
import numpy as np
import pytest
import torch

from oracle import normal_ops as no
from oracle import toeplitz as ot
from oracle.ndft import rel_l2
from oracle.solvers import sense_adjoint, sense_forward
from tests.gpu_util import crandn, cuda, host, rand_coords, report

pytestmark = pytest.mark.gpu

TOL = 2e-5


def problem(shape, C, L, M=2500, seed=0):
    rng = np.random.default_rng(seed)
    coords = rand_coords(rng, len(shape), M)
    w = [rng.uniform(0.2, 1.5, M) * np.exp(1j * rng.uniform(-1, 1, M)) for _ in range(L)]
    K = np.stack([ot.toeplitz_kernel(coords, shape, w[l]) for l in range(L)]).astype(np.complex64)
    R = np.exp(1j * rng.uniform(-np.pi, np.pi, (L,) + shape)).astype(np.complex64)
    S = crandn(rng, (C,) + shape)
    x = crandn(rng, shape)
    return rng, coords, K, R, S, x


CASES = [((16, 16, 16), 4, 3, 0), ((12, 10, 8), 5, 2, 2), ((32, 24), 3, 3, 0)]


@pytest.mark.parametrize("shape,C,L,chunk", CASES)
def test_idt_t1(shape, C, L, chunk):
    from hastycucompute_amd import mri
    _, _, K, R, S, x = problem(shape, C, L)
    op = mri.NORMAL_IDT_T1_OP(cuda(K), cuda(R), cuda(S), coil_batch=chunk)
    got = host(op(cuda(x)))
    report(f"NORMAL_IDT_T1_OP {shape} C={C} L={L} chunk={chunk}", rel_l2(got, no.normal_idt_t1(x, S, K, R)), TOL)


@pytest.mark.parametrize("shape,C,L,chunk", CASES)
def test_idtw_t1(shape, C, L, chunk):
    from hastycucompute_amd import mri
    rng, _, K, R, S, x = problem(shape, C, L, seed=1)
    A = crandn(rng, (C, C)).astype(np.complex128)
    Wm = (A @ A.conj().T / C + np.eye(C)).astype(np.complex64)
    op = mri.NORMAL_IDTW_T1_OP(cuda(K), cuda(R), cuda(S), cuda(Wm), coil_batch=chunk)
    got = host(op(cuda(x)))
    report(f"NORMAL_IDTW_T1_OP {shape} C={C} L={L}", rel_l2(got, no.normal_idtw_t1(x, S, K, R, Wm)), TOL)


@pytest.mark.parametrize("shape,C,L,chunk", CASES)
def test_t2_family(shape, C, L, chunk):
    from hastycucompute_amd import mri
    rng, _, K, R, _, x = problem(shape, C, L, seed=2)
    psi = crandn(rng, (C,) + shape)
    Wm = crandn(rng, (C, C))
    kx, kR, kK, kpsi = cuda(x), cuda(R), cuda(K), cuda(psi)
    got = host(mri.NORMAL_T_T2_OP(kK[0].contiguous(), kx, stack_batch=chunk)(kpsi))
    report(f"NORMAL_T_T2_OP {shape}", rel_l2(got, no.normal_t_t2(psi, x, K[0])), TOL)
    got = host(mri.NORMAL_IDT_T2_OP(kK, kR, kx, stack_batch=chunk)(kpsi))
    report(f"NORMAL_IDT_T2_OP {shape} L={L}", rel_l2(got, no.normal_idt_t2(psi, x, K, R)), TOL)
    got = host(mri.NORMAL_IDTW_T2_OP(kK, kR, kx, cuda(Wm), stack_batch=chunk)(kpsi))
    report(f"NORMAL_IDTW_T2_OP {shape} L={L}", rel_l2(got, no.normal_idtw_t2(psi, x, K, R, Wm)), TOL)


def test_t2_op_changes_stack_size_and_repeats():
    from hastycucompute_amd import mri
    rng, _, K, R, _, x = problem((10, 12, 8), 3, 2, seed=3)
    op = mri.NORMAL_IDT_T2_OP(cuda(K), cuda(R), cuda(x))
    for B in (3, 1, 3):
        psi = crandn(rng, (B,) + x.shape)
        report(f"NORMAL_IDT_T2_OP B={B}", rel_l2(host(op(cuda(psi))), no.normal_idt_t2(psi, x, K, R)), TOL)


def test_idt_t1_single_kernel_matches_normal_t_t1():
    """L = 1 with unit diagonals is NORMAL_T_T1_OP."""
    from hastycucompute_amd import mri
    rng, _, K, _, S, x = problem((16, 12, 10), 4, 1, seed=4)
    ones = np.ones((1,) + x.shape, np.complex64)
    a = host(mri.NORMAL_IDT_T1_OP(cuda(K), cuda(ones), cuda(S))(cuda(x)))
    b = host(mri.NORMAL_T_T1_OP(cuda(K[0]), cuda(S))(cuda(x)))
    report("IDT_T1(L=1) vs T_T1", rel_l2(a, b), 1e-6)


def test_kernel_count_mismatch_raises_on_call():
    from hastycucompute_amd import mri
    _, _, K, R, S, x = problem((8, 8, 8), 2, 2, M=300, seed=5)
    op = mri.NORMAL_IDT_T1_OP(cuda(K), cuda(R[:1]), cuda(S))
    with pytest.raises(RuntimeError, match="Number of kernels and kernel diagonals must match."):
        op(cuda(x))


def test_kernels_built_on_gpu():
    """Kernels from hasty_toeplitz_kernel with weights, end to end."""
    from hastycucompute_amd import fft as hf, mri
    rng = np.random.default_rng(6)
    shape, M, L, C = (12, 12, 12), 3000, 2, 3
    coords = rand_coords(rng, 3, M)
    w = [rng.uniform(0.5, 1.0, M).astype(np.complex64) for _ in range(L)]
    kg = torch.empty((L,) + tuple(2 * n for n in shape), dtype=torch.complex64, device="cuda")
    for l in range(L):
        k = torch.empty(tuple(2 * n for n in shape), dtype=torch.complex64, device="cuda")
        hf.toeplitz_kernel([cuda(c) for c in coords], k, cuda(w[l].reshape(1, -1)))
        kg[l].copy_(k)
    R = np.exp(1j * rng.uniform(-1, 1, (L,) + shape)).astype(np.complex64)
    S = crandn(rng, (C,) + shape)
    x = crandn(rng, shape)
    K = np.stack([ot.toeplitz_kernel(coords, shape, w[l]) for l in range(L)])
    got = host(mri.NORMAL_IDT_T1_OP(kg, cuda(R), cuda(S))(cuda(x)))
    report("IDT_T1 with GPU-built kernels", rel_l2(got, no.normal_idt_t1(x, S, K, R)), TOL)


# ---------------------------------------------------------------------------
# time-segmented SENSE
# ---------------------------------------------------------------------------
def offres_problem(shape, C, L, M, seed, dtype=np.complex64):
    rng = np.random.default_rng(seed)
    coords = rand_coords(rng, len(shape), M)
    # time segments: b_l(t_j) interpolation weights, c_l = exp(-i w t_l)
    t = rng.uniform(0, 1, M)
    tl = np.linspace(0, 1, L)
    bil = np.maximum(0, 1 - np.abs(t[None] - tl[:, None]) * (L - 1)) if L > 1 else np.ones((1, M))
    fmap = rng.uniform(-3, 3, shape)
    cjl = np.exp(-1j * fmap[None] * tl.reshape((L,) + (1,) * len(shape)))
    S = crandn(rng, (C,) + shape, dtype)
    x = crandn(rng, shape, dtype)
    return rng, coords, bil.astype(dtype), cjl.astype(dtype), S, x


@pytest.mark.parametrize("shape,C,L,chunk", [((16, 16, 16), 4, 3, 0), ((16, 12, 10), 5, 2, 2),
                                             ((40, 32), 3, 4, 0)])
def test_offresonance_forward_adjoint(shape, C, L, chunk):
    from hastycucompute_amd import mri
    rng, coords, bil, cjl, S, x = offres_problem(shape, C, L, 4000, seed=7)
    op = mri.SenseOffresonance([cuda(c) for c in coords], cuda(S), cuda(bil), cuda(cjl), coil_batch=chunk)
    y = host(op.forward(cuda(x)))
    report(f"sense_forward_offres {shape} C={C} L={L} chunk={chunk}",
           rel_l2(y, no.sense_forward_offres(x, S, coords, bil, cjl)), TOL)
    g = crandn(rng, (C, len(coords[0])))
    ref = no.sense_adjoint_offres(g, S, coords, bil, cjl)
    xa = host(op.adjoint(cuda(g)))
    report(f"sense_adjoint_offres {shape}", rel_l2(xa, ref), TOL)
    xc = host(op.adjoint(cuda(g), combine=True))
    report(f"sense_adjoint_offres combined {shape}", rel_l2(xc, ref.sum(0)), TOL)


def test_offresonance_functions_and_f64():
    from hastycucompute_amd import mri
    shape, C, L = (12, 10, 8), 3, 3
    rng, coords, bil, cjl, S, x = offres_problem(shape, C, L, 2000, seed=8, dtype=np.complex128)
    c64 = [torch.from_numpy(c.astype(np.float64)).cuda() for c in coords]
    y = host(mri.sense_forward_offresonance(c64, cuda(x), cuda(S), cuda(bil), cuda(cjl)))
    report("sense_forward_offresonance f64", rel_l2(y, no.sense_forward_offres(x, S, coords, bil, cjl)), 1e-11)
    g = crandn(rng, (C, len(coords[0])), np.complex128)
    xa = host(mri.sense_adjoint_offresonance(c64, cuda(g), cuda(S), cuda(bil), cuda(cjl)))
    report("sense_adjoint_offresonance f64", rel_l2(xa, no.sense_adjoint_offres(g, S, coords, bil, cjl)), 1e-11)


def test_sense_nufft_fused_l1():
    from hastycucompute_amd import mri
    rng = np.random.default_rng(9)
    shape, C, M = (16, 20, 12), 4, 5000
    coords = rand_coords(rng, 3, M)
    S = crandn(rng, (C,) + shape)
    x = crandn(rng, shape)
    op = mri.SenseNufft([cuda(c) for c in coords], cuda(S), tol=1e-6)
    report("SenseNufft forward", rel_l2(host(op.forward(cuda(x))), sense_forward(x, S, coords)), TOL)
    g = crandn(rng, (C, M))
    report("SenseNufft adjoint", rel_l2(host(op.adjoint(cuda(g))), sense_adjoint(g, S, coords)), TOL)


def test_offresonance_with_fitted_interpolators_vs_exact_model():
    """Field-map correction end to end: interpolators fitted to a smooth rate
    map, the HIP time-segmented forward against the oracle on the same B, Ct
    (parity), and against the exact off-resonance model
    y[c, j] = sum_n S_c x e^{-i k_j.n} e^{z_n t_j} by a dense sum (the fit)."""
    from hastycucompute_amd import mri
    from hastycucompute_amd.offres import make_offresonance_interpolators
    from oracle.ndft import centred_modes
    rng = np.random.default_rng(10)
    shape, C, M, L = (24, 20), 2, 1500, 6
    coords = rand_coords(rng, 2, M)
    yy, xx = np.meshgrid(*[np.linspace(-1, 1, n) for n in shape], indexing="ij")
    fmap = (-15.0 + 1j * 2 * np.pi * 40.0 * (xx ** 2 - 0.4 * yy)).astype(np.complex64)
    t = np.sort(rng.uniform(0, 6e-3, M))
    B, Ct = make_offresonance_interpolators(cuda(fmap), cuda(t), L, 32)
    S = crandn(rng, (C,) + shape)
    x = crandn(rng, shape)
    op = mri.SenseOffresonance([cuda(c) for c in coords], cuda(S), B, Ct)
    y = host(op.forward(cuda(x)))
    report("offres forward (fitted B, Ct) vs oracle", rel_l2(y, no.sense_forward_offres(x, S, coords, host(B), host(Ct))), TOL)
    grids = np.meshgrid(*[centred_modes(n) for n in shape], indexing="ij")
    phase = sum(np.outer(np.asarray(coords[a], np.float64), grids[a].ravel()) for a in range(2))
    E = np.exp(-1j * phase) * np.exp(np.outer(t, fmap.astype(np.complex128).ravel()))
    exact = (S.reshape(C, -1).astype(np.complex128) * x.ravel()[None]) @ E.T
    err = rel_l2(y, exact)
    print(f"[fit] time-segmented forward vs exact off-resonance model, L={L}: {err:.3e}", flush=True)
    assert err < 5e-3
