"""GPU: batched dynamic frames (hastycucompute_amd.frames) and the NUFFT
autotuner (hastycucompute_amd.autotune).

* The hipGraph replay of one CG iteration of every frame gives the same
  result as the eager launches (bitwise: every kernel is deterministic), on
  two solves in a row with eager work between them.
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
    from hastycucompute_amd.frames import FrameBatch
    n, S, coords, kdata = _frames_problem()
    fb = FrameBatch(cuda(S), [[cuda(c) for c in k] for k in coords], [cuda(y) for y in kdata], tol=1e-6)
    iters = 8
    eager = fb.solve(iters, graph=False).clone()
    g1 = fb.solve(iters, graph=True).clone()
    torch.cuda.synchronize()
    _ = (eager * 2).sum().item()            # eager work between the replays
    g2 = fb.solve(iters, graph=True).clone()
    torch.cuda.synchronize()
    assert torch.equal(g1, eager), "graph replay differs from the eager launches"
    assert torch.equal(g2, eager), "second graph solve differs from the eager launches"
    print(f"[parity] frames: graph == eager bitwise over {iters} iterations x {len(coords)} frames", flush=True)
    V = float(np.prod(n))
    for f in range(len(coords)):
        k64 = [c.astype(np.float64) for c in coords[f]]
        K = ot.toeplitz_kernel(k64, n)
        b = sum(np.conj(S[c]) * ndft_type1(kdata[f][c].astype(np.complex128), k64, n)
                for c in range(S.shape[0])) / V
        xo = conjugate_gradient(lambda v: sense_normal(v, S, K), b, np.zeros(n, np.complex128), iters)
        report(f"frame {f} ({len(coords[f][0])} pts) CG vs oracle FP64 CG", rel_l2(host(eager[f]), xo), 1e-4)


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
