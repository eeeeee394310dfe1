"""GPU: the reconstruction workloads end to end.

* SENSE-CG reconstruction at 64^3 x 8 coils, kooshball 1024 x 128, y = A x
  + 1% noise, 20 CG iterations: the right-hand side against the oracle's KB
  adjoint, the first CG iterates against the oracle's FP64 CG on the same
  Toeplitz kernel, and after 20 iterations the residual and the error vs the
  truth against FP64 CG's (the iterates themselves drift apart in FP32 on
  this ill-conditioned system; the difference is printed).
* 128^3 x 32 coils x 8.4M points (the benchmark size): type 1, type 2 and
  A^H A against FP64 references (direct sums, complex128 plans), a
  30-iteration CG; and through properties
  that hold at any size: NUFFT adjointness <A f, c> = <f, A^H c>, the normal
  operator Hermitian and positive semi-definite, and the Toeplitz normal
  operator equal to (1/N^3) sum_c conj(S_c) A^H A S_c from the NUFFT pair.
* Trajectory design at 256^2 x 8 coils, 16 spiral interleaves x 4096
  samples: 200 Adam steps on the trajectory lower the data loss.
* Gradient through the SENSE-CG solve at 160^3 x 16 coils, stack of spirals
  40 x 24 x 2048 (1.97M points), 25 CG iterations: the reverse-mode adjoint
  through the checkpointed iterates of the FP64 pipeline (complex128 plans,
  FP64 CG) agrees with FP64 central differences along 8 random trajectory
  directions to 1e-4; the production FP32 pipeline (timed) agrees with the
  same differences to 1e-3.
"""
import numpy as np
import pytest
import torch

from oracle.kb import KBPlan
from oracle.ndft import rel_l2
from oracle.solvers import conjugate_gradient as cg_ref
from oracle.solvers import sense_normal
from tests.gpu_util import host, report

pytestmark = pytest.mark.gpu


def _dot64(a, b):
    return complex(torch.vdot(a.reshape(-1).to(torch.complex128), b.reshape(-1).to(torch.complex128)))


def test_sense_recon_64c8_kooshball_vs_oracle():
    from hastycucompute_amd import fft as hf, synth
    from hastycucompute_amd.mri import NORMAL_T_T1_OP
    from hastycucompute_amd.op import conjugate_gradient
    N, C, spokes, samples = 64, 8, 1024, 128
    n = (N, N, N)
    dev = torch.device("cuda", 0)
    k_np = synth.kooshball(spokes, samples)
    k = [torch.from_numpy(c).to(dev) for c in k_np]
    S_np = synth.coil_maps(n, C, seed=1)
    x_np = synth.image(n, seed=0)
    S, x = torch.from_numpy(S_np).to(dev), torch.from_numpy(x_np).to(dev)
    M = spokes * samples
    p2 = hf.nufft_plan.make(hf.nufft_opts(nmodes=n, ntransf=C, tol=1e-5, sign=-1), hf.nufft_type.TYPE_2)
    p1 = hf.nufft_plan.make(hf.nufft_opts(nmodes=n, ntransf=C, tol=1e-5, sign=1), hf.nufft_type.TYPE_1)
    p2.setpts(k)
    p1.setpts(k)
    y = torch.empty((C, M), dtype=torch.complex64, device=dev)
    p2.execute((S * x[None]).contiguous(), y)
    g = torch.Generator(device=dev).manual_seed(2)
    noise = torch.randn((C, M), dtype=torch.complex64, device=dev, generator=g)
    y = (y + 0.01 * y.norm() / noise.norm() * noise).contiguous()
    back = torch.empty((C,) + n, dtype=torch.complex64, device=dev)
    p1.execute(y, back)
    V = float(N ** 3)
    b = ((S.conj() * back).sum(0) / V).contiguous()
    # right-hand side vs the oracle's KB adjoint (FP64, same width)
    plan = KBPlan(n, 1e-5).setpts([c.astype(np.float64) for c in k_np])
    back_ref = plan.type1(host(y).astype(np.complex128))
    b_ref = (np.conj(S_np) * back_ref).sum(0) / V
    report("64^3 x 8 right-hand side vs oracle KB adjoint", rel_l2(host(b), b_ref), 2e-6)
    kernel = torch.empty((2 * N,) * 3, dtype=torch.complex64, device=dev)
    hf.toeplitz_kernel(k, kernel)
    A = NORMAL_T_T1_OP(kernel, S)
    K = host(kernel).astype(np.complex128)
    Nref = lambda v: sense_normal(v, S_np, K, workers=8)   # noqa: E731
    bh = host(b)
    # the first iterations follow the FP64 recurrence closely ...
    for it, tol in ((1, 2e-6), (3, 2e-5)):
        xh = torch.zeros_like(x)
        conjugate_gradient(A, it, 0.0)(xh, b)
        report(f"64^3 x 8 CG {it} iterations vs oracle FP64 CG",
               rel_l2(host(xh), cg_ref(Nref, bh, np.zeros(n), it)), tol)
    # ... and 20 iterations of this undersampled (ill-conditioned) system
    # reach the same residual and error vs the truth as FP64 CG, although
    # FP32 rounding steers the iterate along a different path
    xh = torch.zeros_like(x)
    conjugate_gradient(A, 20, 0.0)(xh, b)
    x_ref = cg_ref(Nref, bh, np.zeros(n), 20)
    res_gpu = np.linalg.norm(Nref(host(xh)) - bh) / np.linalg.norm(bh)
    res_ref = np.linalg.norm(Nref(x_ref) - bh) / np.linalg.norm(bh)
    err_gpu, err_ref = rel_l2(host(xh), x_np), rel_l2(x_ref, x_np)
    print(f"[workload] 64^3 x 8, 20 CG iterations: residual {res_gpu:.4e} (FP64 CG {res_ref:.4e}), "
          f"error vs truth {err_gpu:.4e} (FP64 CG {err_ref:.4e}), iterate difference "
          f"{rel_l2(host(xh), x_ref):.3e}", flush=True)
    assert res_gpu <= 1.1 * res_ref and abs(err_gpu - err_ref) <= 0.02 * err_ref


def test_full_size_properties_128c32_koosh8M():
    from hastycucompute_amd import fft as hf, synth
    from hastycucompute_amd.mri import NORMAL_T_T1_OP
    N, C, spokes, samples = 128, 32, 32768, 256
    n = (N, N, N)
    dev = torch.device("cuda", 0)
    k = [torch.from_numpy(c).to(dev) for c in synth.kooshball(spokes, samples)]
    M = spokes * samples
    S = torch.from_numpy(synth.coil_maps(n, C, seed=1)).to(dev)
    g = torch.Generator(device=dev).manual_seed(7)
    f = torch.randn((C,) + n, dtype=torch.complex64, device=dev, generator=g)
    c = torch.randn((C, M), dtype=torch.complex64, device=dev, generator=g)
    p2 = hf.nufft_plan.make(hf.nufft_opts(nmodes=n, ntransf=C, tol=1e-5, sign=-1), hf.nufft_type.TYPE_2)
    p1 = hf.nufft_plan.make(hf.nufft_opts(nmodes=n, ntransf=C, tol=1e-5, sign=1), hf.nufft_type.TYPE_1)
    p2.setpts(k)
    p1.setpts(k)
    Af = torch.empty((C, M), dtype=torch.complex64, device=dev)
    p2.execute(f, Af)
    AHc = torch.empty((C,) + n, dtype=torch.complex64, device=dev)
    p1.execute(c, AHc)
    lhs, rhs = _dot64(Af, c), _dot64(f, AHc)
    report("full size <A f, c> vs <f, A^H c>", abs(lhs - rhs) / abs(lhs), 2e-5)
    kernel = torch.empty((2 * N,) * 3, dtype=torch.complex64, device=dev)
    hf.toeplitz_kernel(k, kernel)
    op = NORMAL_T_T1_OP(kernel, S)
    u = torch.randn(n, dtype=torch.complex64, device=dev, generator=g)
    v = torch.randn(n, dtype=torch.complex64, device=dev, generator=g)
    Nu, Nv = op(u), op(v)
    a, b = _dot64(u, Nv), _dot64(Nu, v)
    report("full size normal operator Hermitian <u, N v> vs <N u, v>", abs(a - b) / abs(a), 1e-5)
    q = _dot64(v, Nv)
    assert q.real > 0 and abs(q.imag) <= 1e-5 * q.real, q
    print(f"[parity] full size <v, N v> = {q.real:.6e} (imag/real {abs(q.imag) / q.real:.1e})", flush=True)
    # Toeplitz normal operator vs the NUFFT pair
    Sv = (S * v[None]).contiguous()
    p2.execute(Sv, Af)
    p1.execute(Af, AHc)
    pair = (S.conj() * AHc).sum(0) / float(N ** 3)
    report("full size Toeplitz normal op vs NUFFT pair", rel_l2(host(Nv), host(pair)), 2e-5)


def test_full_size_fp64_accuracy_128c32_koosh8M():
    """The accuracy contract at the benchmark size (128^3 x 32 coils x 8.4M
    points, eps 1e-5), against FP64 references:

    * type 1 of every coil's data on 256 random voxels (coils 0, 13 and 31)
      against a direct FP64 sum over all 8.4M points;
    * type 2 on 512 random points of coils 0 and 31 against the FP64 NDFT;
    * the whole A^H A apply (Toeplitz, c64) against complex128 NUFFT plans at
      tol 1e-9, (1/N^3) sum_c conj(S_c) A1(A2(S_c x));
    * 30 CG iterations on N x = b (FP64 scalars): the recurrence residual
      agrees with the true residual ||b - N x|| of a fresh apply, and the
      residual falls (both printed).
    """
    import bench
    from hastycucompute_amd import fft as hf, synth
    from hastycucompute_amd.mri import NORMAL_T_T1_OP
    from hastycucompute_amd.op import conjugate_gradient
    N, C, spokes, samples = 128, 32, 32768, 256
    n = (N, N, N)
    dev = torch.device("cuda", 0)
    k = [torch.from_numpy(c).to(dev) for c in synth.kooshball(spokes, samples)]
    M = spokes * samples
    S = torch.from_numpy(synth.coil_maps(n, C, seed=1)).to(dev)
    x = torch.from_numpy(synth.image(n, seed=0)).to(dev)
    g = torch.Generator(device=dev).manual_seed(11)
    c = torch.randn((C, M), dtype=torch.complex64, device=dev, generator=g)
    p1 = hf.nufft_plan.make(hf.nufft_opts(nmodes=n, ntransf=C, tol=1e-5), hf.nufft_type.TYPE_1)
    p1.setpts(k)
    img = torch.empty((C,) + n, dtype=torch.complex64, device=dev)
    p1.execute(c, img)
    p1.free()
    gc = torch.Generator().manual_seed(3)
    vox = torch.stack([torch.randint(-(N // 2), (N + 1) // 2, (256,), generator=gc) for _ in range(3)], 1).to(dev)
    vidx = tuple(vox[:, a] + N // 2 for a in range(3))
    for cc in (0, 13, C - 1):
        ref = bench.fp64_direct_type1_subset(c[cc], k, n, vox)
        report(f"128^3 x 32 type 1 coil {cc} vs FP64 direct sum (256 voxels)",
               float((img[cc][vidx].to(torch.complex128) - ref).norm() / ref.norm()), 2e-5)
    del c
    p2 = hf.nufft_plan.make(hf.nufft_opts(nmodes=n, ntransf=C, tol=1e-5), hf.nufft_type.TYPE_2)
    p2.setpts(k)
    f = (S * x[None]).contiguous()
    y = torch.empty((C, M), dtype=torch.complex64, device=dev)
    p2.execute(f, y)
    p2.free()
    idx = torch.randperm(M, generator=gc)[:512].to(dev)
    for cc in (0, C - 1):
        ref = bench.fp64_ndft_type2_subset(f[cc], k, idx)
        report(f"128^3 x 32 type 2 coil {cc} vs FP64 NDFT (512 points)",
               float((y[cc][idx].to(torch.complex128) - ref).norm() / ref.norm()), 2e-5)
    del y, f, img
    torch.cuda.empty_cache()
    kernel = torch.empty((2 * N,) * 3, dtype=torch.complex64, device=dev)
    hf.toeplitz_kernel(k, kernel)
    op = NORMAL_T_T1_OP(kernel, S)
    out = op(x)
    ref_n = bench.fp64_plan_pair_normal(k, S, x, N, tol=1e-9)
    report("128^3 x 32 A^H A vs complex128 NUFFT plans (tol 1e-9)",
           float((out.to(torch.complex128) - ref_n).norm() / ref_n.norm()), 2e-5)
    del ref_n
    torch.cuda.empty_cache()
    b = op(x)
    xs = torch.zeros_like(x)
    hist = []
    conjugate_gradient(op, 30, 0.0)(xs, b, history=hist)
    bn = float(b.to(torch.complex128).norm())
    rec = float(hist[-1].item()) / bn
    true = float((b - op(xs)).to(torch.complex128).norm()) / bn
    print(f"[workload] 128^3 x 32, 30 CG iterations: residual {rec:.4e} (recurrence), {true:.4e} (true), "
          f"first {float(hist[0].item()) / bn:.4e}", flush=True)
    report("128^3 x 32 CG recurrence residual vs true residual", abs(rec - true) / true, 1e-3)
    assert rec < float(hist[0].item()) / bn


def test_trajectory_design_spiral_256c8():
    from hastycucompute_amd import synth, trajectory
    dev = torch.device("cuda", 0)
    n, C, I, Sm = (256, 256), 8, 16, 4096
    x = torch.from_numpy(synth.ellipse_phantom(n, seed=4)).to(dev)
    S = torch.from_numpy(synth.coil_maps(n, C, seed=5)).to(dev)
    k_true = [torch.from_numpy(c).to(dev) for c in synth.spiral2d(I, Sm)]
    from hastycucompute_amd.autograd import NufftOperator, sense_forward
    with torch.no_grad():
        y = sense_forward(x, S, NufftOperator(n, ntransf=C, tol=1e-5), *k_true)
    g = torch.Generator(device=dev).manual_seed(6)
    k0 = [kt + 0.02 * torch.randn(kt.shape, device=dev, generator=g) for kt in k_true]
    out = trajectory.design(x, S, k0, y, ninterleaves=I, steps=200, lr=2e-3, gmax=0.2)
    d0, d1 = out["data_loss"][0], out["data_loss"][-1]
    print(f"[workload] spiral design 256^2 x 8 coils x 65536 pts: data loss {d0:.4e} -> {d1:.4e} "
          f"in 200 Adam steps", flush=True)
    assert all(np.isfinite(out["loss"]))
    assert d1 < 0.5 * d0


def test_cg_gradient_160c16_stack_of_spirals():
    """cfg5: the gradient of L(k) = ||x_25(k) - x_true||^2 / ||x_true||^2 through
    25 CG iterations, 160^3 x 16 coils, stack of spirals 40 x 24 x 2048.

    * The FP64 pipeline (complex128 NUFFT plans at tol 1e-8, FP64 CG on the
      normal operator of the NUFFT pair): the reverse-mode adjoint through the
      checkpointed iterates against FP64 central differences of the same
      pipeline (h = 1e-2 along unit directions in trajectory space), on 8
      random directions: <= 1e-4 of the RMS directional derivative.
    * The production FP32 pipeline (Toeplitz normal operator at tol 1e-5,
      FP64 CG scalars; timed): its adjoint against the same FP64 central
      differences, <= 1e-3 (measured 3e-4: FP32 rounding in the 25 CG steps
      and the tolerance gap between the two pipelines).
    """
    import time
    from hastycucompute_amd import autograd as ag, synth
    dev = torch.device("cuda", 0)
    N, C, iters = 160, 16, 25
    n = (N, N, N)
    k = [torch.from_numpy(c).to(dev) for c in synth.stack_of_spirals(40, 24, 2048)]
    M = int(k[0].shape[0])
    S = torch.from_numpy(synth.coil_maps(n, C, seed=1)).to(dev)
    x_true = torch.from_numpy(synth.image(n, seed=0)).to(dev)
    xt64 = x_true.to(torch.complex128)
    # ---- FP64 pipeline: adjoint gradient and central differences ----
    op64 = ag.NufftOperator(n, ntransf=C, tol=1e-8, dtype=torch.float64)
    k64 = [c.double() for c in k]
    S64 = S.to(torch.complex128)
    with torch.no_grad():
        y64 = ag.sense_forward(xt64, S64, op64, *k64).contiguous()

    def loss64(kk):
        xh = ag.sense_cg_solve(y64, S64, op64, *kk, iters=iters)
        return (xh - xt64).abs().pow(2).sum() / xt64.abs().pow(2).sum()

    kg64 = [c.clone().requires_grad_(True) for c in k64]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    L64 = loss64(kg64)
    L64.backward()
    torch.cuda.synchronize()
    g64 = [c.grad for c in kg64]
    print(f"[workload] 160^3 x 16 coils x {M} pts FP64 pipeline: loss {float(L64.detach()):.6e}, "
          f"forward + backward {time.perf_counter() - t0:.1f} s", flush=True)
    gen = torch.Generator(device=dev).manual_seed(8)
    dirs, fds = [], []
    h = 1e-2
    for _ in range(8):
        d = [torch.randn(c.shape, device=dev, dtype=torch.float64, generator=gen) for c in k]
        nrm = float(torch.sqrt(sum((di ** 2).sum() for di in d)))
        d = [di / nrm for di in d]
        with torch.no_grad():
            lp = float(loss64([c + h * di for c, di in zip(k64, d)]))
            lm = float(loss64([c - h * di for c, di in zip(k64, d)]))
        dirs.append(d)
        fds.append((lp - lm) / (2 * h))
    an64 = [float(sum((gi * di).sum() for gi, di in zip(g64, d))) for d in dirs]
    scale = float(np.sqrt(np.mean(np.square(fds))))
    for a_, f_ in zip(an64, fds):
        print(f"[workload]   direction: FP64 adjoint {a_:+.9e}  FP64 central difference {f_:+.9e}", flush=True)
    report("160^3 FP64 truncated-CG adjoint vs FP64 central differences (8 directions, max |err| / rms)",
           max(abs(a_ - f_) for a_, f_ in zip(an64, fds)) / scale, 1e-4)
    del op64, kg64, g64, y64
    torch.cuda.empty_cache()
    # ---- FP32 pipeline (production): timed, against the FP64 differences ----
    op = ag.NufftOperator(n, ntransf=C, tol=1e-5)
    with torch.no_grad():
        y = ag.sense_forward(x_true, S, op, *k).contiguous()
    kg = [c.clone().requires_grad_(True) for c in k]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xh = ag.sense_cg_solve(y, S, op, *kg, iters=iters)
    L = (xh.to(torch.complex128) - xt64).abs().pow(2).sum() / xt64.abs().pow(2).sum()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    L.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    g = [c.grad.double() for c in kg]
    assert all(torch.isfinite(gi).all() for gi in g)
    print(f"[workload] 160^3 x 16 coils x {M} pts FP32 pipeline, {iters} CG iterations: forward {t1 - t0:.2f} s "
          f"(incl. PSF), adjoint backward {t2 - t1:.2f} s, loss {float(L.detach()):.6e}", flush=True)
    an32 = [float(sum((gi * di).sum() for gi, di in zip(g, d))) for d in dirs]
    for a_, f_ in zip(an32, fds):
        print(f"[workload]   direction: FP32 adjoint {a_:+.6e}  FP64 central difference {f_:+.6e}", flush=True)
    report("160^3 FP32 truncated-CG adjoint vs FP64 central differences (8 directions, max |err| / rms)",
           max(abs(a_ - f_) for a_, f_ in zip(an32, fds)) / scale, 1e-3)
