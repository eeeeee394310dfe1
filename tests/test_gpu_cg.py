"""GPU parity: FP64 reductions, CG and power iteration vs the oracle."""
import numpy as np
import pytest
import torch

from oracle.ndft import rel_l2
from oracle.solvers import conjugate_gradient as cg_ref
from oracle.solvers import max_eig as max_eig_ref
from oracle.solvers import sense_normal
from tests.gpu_util import crandn, cuda, host, rand_coords, report

pytestmark = pytest.mark.gpu


def test_dot_and_norm_fp64():
    from hastycucompute_amd._lib import check, lib
    rng = np.random.default_rng(1)
    for n in (1, 255, 1000, 1 << 20):
        a = crandn(rng, n)
        b = crandn(rng, n)
        out = torch.zeros(2, dtype=torch.float64, device="cuda")
        ga, gb = cuda(a), cuda(b)   # keep both alive while the library reads them
        check(lib().hasty_dot(ga.data_ptr(), gb.data_ptr(), n, out.data_ptr(), 0))
        ref = np.vdot(a.astype(np.complex128), b.astype(np.complex128))
        got = complex(*host(out))
        assert abs(got - ref) <= 1e-12 * np.sqrt(n) * np.linalg.norm(a) * np.linalg.norm(b), (n, got, ref)
        check(lib().hasty_norm(ga.data_ptr(), n, out.data_ptr(), 0))
        assert abs(host(out)[0] - np.linalg.norm(a.astype(np.complex128))) <= 1e-12 * np.linalg.norm(a)


def test_cg_normal_op_matches_oracle():
    from hastycucompute_amd import fft as hf
    from hastycucompute_amd.mri import NORMAL_T_T1_OP
    from hastycucompute_amd.op import conjugate_gradient
    rng = np.random.default_rng(2)
    shape, C, M = (12, 10, 15), 4, 5000
    coords = rand_coords(rng, 3, M)
    S = crandn(rng, (C,) + shape)
    b = crandn(rng, shape)
    k = torch.empty(tuple(2 * n for n in shape), dtype=torch.complex64, device="cuda")
    hf.toeplitz_kernel([cuda(c) for c in coords], k)
    op = NORMAL_T_T1_OP(k, cuda(S))
    K = host(k).astype(np.complex128)
    for iters in (1, 5, 12):
        x = torch.zeros(shape, dtype=torch.complex64, device="cuda")
        conjugate_gradient(op, iters, 0.0)(x, cuda(b))
        ref = cg_ref(lambda v: sense_normal(v, S, K), b, np.zeros(shape), iters)
        report(f"CG {iters} iterations vs oracle CG", rel_l2(host(x), ref), 2e-4 if iters == 12 else 5e-5)


def test_cg_exact_convergence_small_spd():
    """A diagonal SPD operator with 3 distinct eigenvalues: CG converges in 3 steps."""
    from hastycucompute_amd.op import conjugate_gradient
    n = 1000
    d = np.where(np.arange(n) % 3 == 0, 1.0, np.where(np.arange(n) % 3 == 1, 4.0, 9.0)).astype(np.float32)
    dg = cuda(d.astype(np.complex64))
    rng = np.random.default_rng(3)
    b = crandn(rng, n)
    x = torch.zeros(n, dtype=torch.complex64, device="cuda")
    conjugate_gradient(lambda v: dg * v, 3, 0.0)(x, cuda(b))
    report("CG exact in 3 steps", rel_l2(host(x), b / d), 1e-5)


def test_cg_tolerance_stops_on_device():
    from hastycucompute_amd.op import conjugate_gradient
    n = 512
    rng = np.random.default_rng(4)
    b = crandn(rng, n)
    x = torch.zeros(n, dtype=torch.complex64, device="cuda")
    conjugate_gradient(lambda v: 2.0 * v, 50, 1e-3)(x, cuda(b))
    report("CG stops at tolerance (identity*2)", rel_l2(host(x), b / 2), 1e-6)


def test_max_eig_matches_oracle():
    from hastycucompute_amd.op import max_eig
    n = 300
    rng = np.random.default_rng(5)
    d = np.linspace(0.1, 3.0, n).astype(np.float32)
    dg = cuda(d.astype(np.complex64))
    x0 = crandn(rng, n)
    lam, x = max_eig(lambda v: dg * v, cuda(x0), 30)
    lam_ref, _ = max_eig_ref(lambda v: d * v, x0, 30)
    assert abs(lam - lam_ref) <= 1e-5 * lam_ref, (lam, lam_ref)


def _spd(rng, n, cond=30.0):
    Q = np.linalg.qr(crandn(rng, (n, n)))[0]
    A = Q @ np.diag(np.linspace(1, 6, n)) @ Q.conj().T + np.diag(np.linspace(1, cond, n))
    return A.astype(np.complex64).astype(np.complex128)     # the matrix the GPU applies


def test_pcg_jacobi_matches_oracle_and_dense_solve():
    """Preconditioned CG (``cg.cppm:16-119``) with a Jacobi preconditioner on a
    dense SPD matrix applied in complex64: early iterates against the oracle's
    FP64 PCG, the solution after n steps against numpy.linalg.solve."""
    from hastycucompute_amd.op import conjugate_gradient
    from oracle.solvers import preconditioned_conjugate_gradient as pcg_ref
    rng = np.random.default_rng(30)
    n = 48
    A = _spd(rng, n)
    J = (1.0 / np.diag(A).real).astype(np.float32)
    gA, gJ = cuda(A.astype(np.complex64)), cuda(J.astype(np.complex64))
    Af = lambda v: torch.matmul(gA, v)
    Pf = lambda r: gJ * r
    b = crandn(rng, n)
    for k in (1, 3, 6):
        x = torch.zeros(n, dtype=torch.complex64, device="cuda")
        conjugate_gradient(Af, Pf, k, 1, 0.0)(x, cuda(b))
        ref = pcg_ref(lambda v: A @ v, lambda r: J * r, b, np.zeros(n), k, 1, 0.0)
        report(f"PCG (Jacobi) {k} steps vs oracle FP64 PCG", rel_l2(host(x), ref), 2e-5)
    x = torch.zeros(n, dtype=torch.complex64, device="cuda")
    conjugate_gradient(Af, Pf, 2 * n, 1, 0.0)(x, cuda(b))
    report("PCG (Jacobi) 2n steps vs dense FP64 solve", rel_l2(host(x), np.linalg.solve(A, b)), 1e-4)


def test_pcg_inverse_preconditioner_converges_in_one_step():
    """With P = A^-1 the first direction is A^-1 r0 and alpha = 1, so one step
    reaches the solution; the device flag then stops the rest."""
    from hastycucompute_amd.op import conjugate_gradient
    rng = np.random.default_rng(31)
    n = 40
    A = _spd(rng, n, cond=5.0)
    gA = cuda(A.astype(np.complex64))
    gAi = cuda(np.linalg.inv(A).astype(np.complex64))
    b = crandn(rng, n)
    x = torch.zeros(n, dtype=torch.complex64, device="cuda")
    conjugate_gradient(lambda v: torch.matmul(gA, v), lambda r: torch.matmul(gAi, r), 1, 1, 0.0)(x, cuda(b))
    report("PCG with P = A^-1, one step vs dense solve", rel_l2(host(x), np.linalg.solve(A, b)), 2e-5)


def test_pcg_restart_callbacks_match_oracle():
    """should_restart_callback(x) true on the 3rd and 7th step resets p = z and
    ends that outer iteration; after_restart_callback runs once per outer
    iteration; the iterate matches the oracle run with the same schedule."""
    from hastycucompute_amd.op import conjugate_gradient
    from oracle.solvers import preconditioned_conjugate_gradient as pcg_ref
    rng = np.random.default_rng(32)
    n = 64
    A = _spd(rng, n)
    J = (1.0 / np.diag(A).real)
    gA, gJ = cuda(A.astype(np.complex64)), cuda(J.astype(np.complex64))
    b = crandn(rng, n)

    def schedule():
        step = [0]

        def should_restart(_x):
            step[0] += 1
            return step[0] in (3, 7)
        return should_restart
    after = []
    x = torch.zeros(n, dtype=torch.complex64, device="cuda")
    conjugate_gradient(lambda v: torch.matmul(gA, v), lambda r: gJ * r, 5, 3, 0.0)(
        x, cuda(b), schedule(), lambda _x: after.append(1))
    ref = pcg_ref(lambda v: A @ v, lambda r: J * r, b, np.zeros(n), 5, 3, 0.0, should_restart=schedule())
    assert len(after) == 3
    report("PCG with restarts at steps 3 and 7 vs oracle", rel_l2(host(x), ref), 5e-5)
