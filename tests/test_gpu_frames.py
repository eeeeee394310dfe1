"""GPU: batched dynamic frames (hastycucompute_amd.frames) and the NUFFT
autotuner (hastycucompute_amd.autotune).

* The hipGraph replay of the whole solve (right-hand sides and CG of every
  frame, as concurrent branches or as one batched system of frames) gives
  the same result as the eager per-frame launches (bitwise: every kernel is
  deterministic), on three replays with eager work between them; at full
  size (20 frames of 96^3 x 16 coils) too, with each frame's normal operator
  Hermitian and positive semi-definite.
* The batched frames operator (one launch per Toeplitz pass over frames x
  coils) equals each frame's own normal operator bitwise, and the oracle.
* Each frame's CG iterate matches the oracle's FP64 CG on the oracle's
  Toeplitz normal operator (same iteration count) to 1e-4.
* A tuned plan computes the same transform as the default plan.
"""
import numpy as np
import pytest
import torch

from oracle import toeplitz as ot
from oracle.ndft import ndft_type1, ndft_type2, rel_l2
from oracle.solvers import conjugate_gradient, sense_normal
from tests.gpu_util import crandn, cuda, host, report

pytestmark = pytest.mark.gpu


def _frames_problem(n=(16, 16, 16), C=3, spokes=(40, 70, 100), samples=24):
    from hastycucompute_amd import synth
    rng = np.random.default_rng(21)
    S = synth.coil_maps(n, C, seed=2)
    x = synth.image(n, seed=3)
    coords, kdata = [], []
    for f, ns in enumerate(spokes):
        k = synth.rotated_kooshball(ns, samples, f)
        k64 = [c.astype(np.float64) for c in k]
        y = np.stack([ndft_type2((S[c] * x).astype(np.complex128), k64) for c in range(C)])
        y = y + 0.01 * np.linalg.norm(y) / np.sqrt(y.size) * crandn(rng, y.shape, np.complex128)
        coords.append(k)
        kdata.append(y.astype(np.complex64))
    return n, S, coords, kdata


def test_frames_graph_equals_eager_and_oracle_cg():
    """The solve (right-hand sides + CG) replayed from one graph of concurrent
    branches equals the eager launches bitwise -- one frame after another on
    one stream, and on the branches -- over three replays with eager work
    between them, and every frame matches the oracle's FP64 CG."""
    from hastycucompute_amd.frames import FrameBatch
    n, S, coords, kdata = _frames_problem()
    fb = FrameBatch(cuda(S), [[cuda(c) for c in k] for k in coords], [cuda(y) for y in kdata], tol=1e-6,
                    streams=2)
    iters = 8
    eager = fb.solve(iters, graph=False, branches=False, batched=False).clone()
    eager_br = fb.solve(iters, graph=False, branches=True, batched=False).clone()
    assert torch.equal(eager_br, eager), "eager branches differ from one stream"
    eager_b = fb.solve(iters, graph=False, batched=True).clone()
    assert torch.equal(eager_b, eager), "batched frames differ from the per-frame launches"
    for rep in range(3):
        for batched in (False, True):
            g = fb.solve(iters, graph=True, batched=batched).clone()
            torch.cuda.synchronize()
            _ = (eager * 2).sum().item()            # eager work between the replays
            assert torch.equal(g, eager), f"graph replay {rep} (batched={batched}) differs from the eager launches"
    print(f"[parity] frames: graph (per-frame and batched) == eager bitwise over {iters} iterations x "
          f"{len(coords)} frames, 3 replays", flush=True)
    V = float(np.prod(n))
    for f in range(len(coords)):
        k64 = [c.astype(np.float64) for c in coords[f]]
        K = ot.toeplitz_kernel(k64, n)
        b = sum(np.conj(S[c]) * ndft_type1(kdata[f][c].astype(np.complex128), k64, n)
                for c in range(S.shape[0])) / V
        xo = conjugate_gradient(lambda v: sense_normal(v, S, K), b, np.zeros(n, np.complex128), iters)
        report(f"frame {f} ({len(coords[f][0])} pts) CG vs oracle FP64 CG", rel_l2(host(eager[f]), xo), 1e-4)


@pytest.mark.parametrize("N,F,C,coil_batch", [(64, 3, 3, 0), (64, 3, 5, 2), (96, 2, 2, 0), (128, 2, 2, 0),
                                              (20, 3, 3, 0)])
def test_frames_normal_op_batched_passes(N, F, C, coil_batch):
    """FramesNormalOp (every Toeplitz pass one launch over frames x coils on
    the fused 64 / 96 / 128 planes; a 20^3 grid runs the frames one after
    another) equals the frames' own NORMAL_T_T1_OP applies bitwise -- with
    coil chunks too, whose last one-coil chunk is not fused -- and frame 0
    matches the oracle's FP64 normal operator on the same kernel."""
    from hastycucompute_amd import fft as hf, synth
    from hastycucompute_amd.frames import FramesNormalOp
    from hastycucompute_amd.mri import NORMAL_T_T1_OP
    n = (N, N, N)
    print(f"[case] frames normal op N={N} F={F} C={C} coil_batch={coil_batch}", flush=True)
    S = cuda(synth.coil_maps(n, C, seed=4))
    kernels = torch.empty((F,) + tuple(2 * m for m in n), dtype=torch.complex64, device=S.device)
    coords = []
    for f in range(F):
        k = [cuda(c) for c in synth.rotated_kooshball(60 + 40 * f, 2 * N, f)]
        hf.toeplitz_kernel(k, kernels[f])
        coords.append(k)
    rng = np.random.default_rng(7)
    x = cuda(crandn(rng, (F,) + n))
    op = FramesNormalOp(kernels, S, coil_batch=coil_batch)
    out = op(x)
    for f in range(F):
        ref = NORMAL_T_T1_OP(kernels[f], S, coil_batch=coil_batch)(x[f].contiguous())
        assert torch.equal(out[f], ref), f"frame {f} differs from its own NORMAL_T_T1_OP"
    if N <= 64:
        # the oracle's FP64 normal operator on frame 0's kernel (the kernel
        # itself is pinned against the NDFT PSF in test_gpu_toeplitz.py)
        K = host(kernels[0]).astype(np.complex128)
        ref = sense_normal(host(x[0]).astype(np.complex128), host(S).astype(np.complex128), K)
        report(f"frames op N={N} frame 0 vs oracle FP64 normal operator", rel_l2(host(out[0]), ref), 2e-5)


def test_frames_full_size_96c16x20():
    """The dynamic workload at full size: 20 frames of 96^3 x 16 coils, spoke
    counts U[500, 3000] x 192 samples.  The replayed solve equals the eager
    one-frame-at-a-time launches bitwise; each frame's normal operator is
    Hermitian and positive semi-definite (frames 0, 9, 19); the graph of
    concurrent branches is timed against the eager launches (printed)."""
    import time
    from hastycucompute_amd import synth
    from hastycucompute_amd.frames import FrameBatch
    dev = torch.device("cuda", 0)
    N, C, F, samples, iters = 96, 16, 20, 192, 15
    rng = np.random.default_rng(3)
    spokes = rng.integers(500, 3001, F)
    n = (N, N, N)
    S = torch.from_numpy(synth.coil_maps(n, C, seed=1)).to(dev)
    g = torch.Generator(device=dev).manual_seed(5)
    coords, kdata = [], []
    for f, ns in enumerate(spokes):
        coords.append([torch.from_numpy(c).to(dev) for c in synth.rotated_kooshball(int(ns), samples, f)])
        kdata.append(torch.randn((C, int(ns) * samples), dtype=torch.complex64, device=dev, generator=g))
    fb = FrameBatch(S, coords, kdata, tol=1e-5)
    eager = fb.solve(iters, graph=False, branches=False, batched=False).clone()
    times = {}
    for name, kw in (("eager", dict(graph=False, branches=False, batched=False)),
                     ("graph", dict(graph=True, batched=False)),
                     ("batched graph", dict(graph=True, batched=True))):
        fb.solve(iters, **kw)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x = fb.solve(iters, **kw)
        torch.cuda.synchronize()
        times[name] = time.perf_counter() - t0
        assert torch.equal(x, eager), f"{name} solve differs from the eager launches"
    print(f"[workload] frames 20 x 96^3 x 16, {iters} CG iterations + right-hand sides: eager "
          f"{times['eager'] * 1e3:.1f} ms, graph of {fb.nstreams} branches {times['graph'] * 1e3:.1f} ms "
          f"({times['eager'] / times['graph']:.2f}x), batched graph {times['batched graph'] * 1e3:.1f} ms "
          f"({times['eager'] / times['batched graph']:.2f}x), bitwise equal", flush=True)
    for f in (0, 9, 19):
        u = torch.randn(n, dtype=torch.complex64, device=dev, generator=g)
        v = torch.randn(n, dtype=torch.complex64, device=dev, generator=g)
        Nu, Nv = fb.ops[f](u), fb.ops[f](v)
        a = complex(torch.vdot(u.reshape(-1).to(torch.complex128), Nv.reshape(-1).to(torch.complex128)))
        b = complex(torch.vdot(Nu.reshape(-1).to(torch.complex128), v.reshape(-1).to(torch.complex128)))
        report(f"frame {f} normal operator Hermitian <u, N v> vs <N u, v>", abs(a - b) / abs(a), 1e-5)
        q = complex(torch.vdot(v.reshape(-1).to(torch.complex128), Nv.reshape(-1).to(torch.complex128)))
        assert q.real > 0 and abs(q.imag) <= 1e-5 * q.real, (f, q)


def test_autotune_choice_computes_same_transform():
    from hastycucompute_amd import autotune, fft as hf, synth
    n, C = (24, 24, 24), 4
    k = [cuda(c) for c in synth.kooshball(300, 48)]
    M = int(k[0].shape[0])
    for nt in (hf.nufft_type.TYPE_1, hf.nufft_type.TYPE_2):
        best = autotune.tune_nufft(n, C, k, nt, tol=1e-5, reps=2, use_cache=False)
        assert best["trials"] and best["ms"] == min(t["ms"] for t in best["trials"])
        opts = autotune.plan_options(best)
        sign = 1 if nt == hf.nufft_type.TYPE_1 else -1
        outs = []
        rng = np.random.default_rng(4)
        src = cuda(crandn(rng, (C, M))) if nt == hf.nufft_type.TYPE_1 else cuda(crandn(rng, (C,) + n))
        for o in ({}, opts):
            p = hf.nufft_plan.make(hf.nufft_opts(nmodes=n, ntransf=C, tol=1e-5, sign=sign, **o), nt)
            p.setpts(k)
            dst = torch.empty(((C,) + n) if nt == hf.nufft_type.TYPE_1 else (C, M), dtype=torch.complex64,
                              device="cuda")
            p.execute(src, dst)
            outs.append(host(dst))
        print(f"[autotune] type {int(nt)}: best {opts} {best['ms']:.3f} ms of {len(best['trials'])} candidates",
              flush=True)
        report(f"autotuned type {int(nt)} plan vs default plan", rel_l2(outs[1], outs[0]), 1e-6)
