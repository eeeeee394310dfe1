"""Batched dynamic reconstruction on one GPU: frames with their own
(uneven) trajectories, reconstructed together.

Frame f has trajectory k_f (M_f points), its own Toeplitz PSF kernel K_f and
SENSE normal operator N_f x = (1/N^d) sum_c conj(S_c) A_f^H A_f (S_c x)
(``NORMAL_T_T1_OP``, all coils per launch, coil maps shared by the frames),
and its own type-1 plan with its points registered once.  A solve is, per
frame, the right-hand side b_f = (1/N^d) sum_c conj(S_c) A_f^H y_f (one
all-coil type-1 NUFFT with conj(S_c) and the coil sum fused into its last
FFT pass), then ``iters`` CG iterations from x = 0, every CG scalar on the
device.

The frames share nothing, so a solve runs them as ``streams`` independent
branches: frame f on branch f mod streams, each branch on its own HIP stream
(default priority), joined at the end.  At 96^3 x 16 coils one frame's
launches do not fill the GPU (the fused last Toeplitz pass has 288 blocks of
one per CU), and the branches fill it.  ``solve(graph=True)`` captures the
whole solve -- right-hand sides, CG initialisation and every iteration of
every frame, all branches -- once into a HIP graph (``torch.cuda.CUDAGraph``)
and replays it as one launch.  The eager path runs the same launches in the
same order per frame, on one stream (``streams=1``) or on the branches; all
of them give identical results, bit for bit (every kernel is deterministic
and no frame reads another's data).

``solve(batched=True)`` (the default) runs the CG of all frames as one
system of frames: each iteration is one ``FramesNormalOp`` apply (every
Toeplitz pass one launch over frames x coils, each frame's lines with its
own kernel) and one batched CG update (``hasty_cg_step_batched``, one launch
per kernel for all frames).  The right-hand sides stay per frame (each frame
has its own points) on the branches.  Every frame gets the same result, bit
for bit, as on the per-frame launches.
"""
from __future__ import annotations

import ctypes
from typing import Dict, List, Optional, Sequence

import torch

from . import fft as hf
from ._lib import CG_STATE_DOUBLES, check, i64_array, lib
from .mri import NORMAL_T_T1_OP


class FramesNormalOp:
    """F frames of ``NORMAL_T_T1_OP`` (``normal.cppm:43-93``) that share the
    coil maps, frame f with its own Toeplitz kernel:
    out[f] = sum_c conj(S_c) . T_f(S_c . x[f]).

    ``kernels`` ``[F, 2N...]`` and ``smaps`` ``[C, N...]`` are complex64 GPU
    tensors kept resident; ``apply(x, out)`` takes ``[F, N...]``.  On the
    fused 3-D passes (N x N planes, N = 64, 96, 128) every pass is one launch
    over all frames and coils (``hasty_normal_frames_create``); the workspace
    is F times that of one frame's operator.
    """

    def __init__(self, kernels: torch.Tensor, smaps: torch.Tensor, coil_batch: int = 0):
        for name, t in (("kernels", kernels), ("smaps", smaps)):
            hf._require_gpu(t, name)
            if t.dtype != torch.complex64:
                raise RuntimeError(f"{name} must be complex64")
        d = smaps.dim() - 1
        if d not in (2, 3) or kernels.dim() != d + 1:
            raise RuntimeError("FramesNormalOp requires DIM in {2,3}, kernels [F, 2N...] and maps [C, N...]")
        n = [int(v) for v in smaps.shape[1:]]
        if tuple(kernels.shape[1:]) != tuple(2 * m for m in n):
            raise RuntimeError("kernels must have shape [F, 2N...]")
        self.F = int(kernels.shape[0])
        self._kernels = kernels
        self._smaps = smaps
        self._n = n
        self._device = smaps.device
        h = ctypes.c_void_p()
        check(lib().hasty_normal_frames_create(d, i64_array(n), self.F, int(smaps.shape[0]), smaps.data_ptr(),
                                               kernels.data_ptr(), int(coil_batch), hf._stream(self._device),
                                               ctypes.byref(h)))
        self._h = h

    def apply(self, x: torch.Tensor, out: torch.Tensor) -> torch.Tensor:
        hf._require_gpu(x, "x")
        hf._require_gpu(out, "out")
        if list(x.shape) != [self.F] + self._n or x.dtype != torch.complex64:
            raise RuntimeError(f"x must be complex64 of shape {[self.F] + self._n}")
        if out.shape != x.shape or out.dtype != x.dtype:
            raise RuntimeError("out must match x")
        check(lib().hasty_normal_set_stream(self._h, hf._stream(self._device)))
        check(lib().hasty_normal_apply(self._h, x.data_ptr(), out.data_ptr()))
        return out

    def __call__(self, x: torch.Tensor) -> torch.Tensor:
        return self.apply(x, torch.empty_like(x))

    def free(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            check(lib().hasty_normal_destroy(self._h))
        self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class FrameBatch:
    def __init__(self, smaps: torch.Tensor, coords: Sequence[Sequence[torch.Tensor]],
                 kdata: Sequence[torch.Tensor], tol: float = 1e-5, kernel_tol: float = 1e-9,
                 nufft_opts: Optional[Dict] = None, streams: int = 4):
        if smaps.dtype != torch.complex64 or smaps.device.type != "cuda":
            raise RuntimeError("smaps must be a complex64 GPU tensor [C, N...]")
        if len(coords) != len(kdata) or not coords:
            raise RuntimeError("one trajectory and one k-space array per frame")
        if streams < 1:
            raise RuntimeError("streams must be >= 1")
        self.F = len(coords)
        self.C = int(smaps.shape[0])
        self.n = tuple(int(s) for s in smaps.shape[1:])
        self.nvox = 1
        for s in self.n:
            self.nvox *= s
        dev = smaps.device
        self.device = dev
        self.smaps = smaps.contiguous()
        opts = dict(nufft_opts or {})
        # the frames' kernels, stacked [F, 2N...] (self.kernels[f] is a view)
        self.kernel_stack = torch.empty((self.F,) + tuple(2 * s for s in self.n), dtype=torch.complex64,
                                        device=dev)
        self.kernels: List[torch.Tensor] = []
        self.ops: List[NORMAL_T_T1_OP] = []
        self.adj: List[hf.nufft_plan] = []
        self.coords: List[List[torch.Tensor]] = []
        self.kdata: List[torch.Tensor] = []
        for f in range(self.F):
            kf = [c.contiguous() for c in coords[f]]
            y = kdata[f].contiguous()
            if tuple(y.shape) != (self.C, int(kf[0].shape[0])):
                raise RuntimeError(f"kdata[{f}] must have shape [{self.C}, {int(kf[0].shape[0])}]")
            kernel = self.kernel_stack[f]
            hf.toeplitz_kernel(kf, kernel, tol=kernel_tol)
            self.kernels.append(kernel)
            self.ops.append(NORMAL_T_T1_OP(kernel, self.smaps))
            plan = hf.nufft_plan.make(hf.nufft_opts(nmodes=self.n, ntransf=self.C, tol=tol,
                                                    sign=hf.nufft_sign.DEFAULT_TYPE_1, **opts),
                                      hf.nufft_type.TYPE_1)
            plan.setpts(kf)
            self.adj.append(plan)
            self.coords.append(kf)
            self.kdata.append(y)
        self.frames_op = FramesNormalOp(self.kernel_stack, self.smaps)
        self.nstreams = min(int(streams), self.F)
        self.streams = [torch.cuda.Stream(device=dev) for _ in range(self.nstreams)]
        # per frame: right-hand side and CG workspace
        shape = (self.F,) + self.n
        self.b = torch.empty(shape, dtype=torch.complex64, device=dev)
        self.x = torch.zeros(shape, dtype=torch.complex64, device=dev)
        self.r = torch.empty_like(self.x)
        self.p = torch.empty_like(self.x)
        self.Ap = torch.empty_like(self.x)
        self.state = torch.zeros((self.F, CG_STATE_DOUBLES), dtype=torch.float64, device=dev)
        self.graphs: Dict[tuple, torch.cuda.CUDAGraph] = {}
        # the right-hand sides once, eagerly (they also warm every plan-owned buffer)
        for f in range(self.F):
            self._rhs(f)
        torch.cuda.synchronize(dev)

    # -- launches of one frame (capturable: kernels only, no allocation in the library) --
    def _rhs(self, f: int) -> None:
        """b_f = (1/N^d) sum_c conj(S_c) A_f^H y_f: one type-1 NUFFT of all
        coils with conj(S_c) and the coil sum in its last FFT pass."""
        plan = self.adj[f]
        s = torch.cuda.current_stream(self.device).cuda_stream
        check(lib().hasty_nufft_set_stream(plan._handle(), s))
        check(lib().hasty_sense_adjoint(plan._handle(), self.C, 1, self.kdata[f].data_ptr(),
                                        self.smaps.data_ptr(), None, None, self.b[f].data_ptr(), 1))
        self.b[f].mul_(1.0 / float(self.nvox))

    def _init(self, f: int) -> None:
        s = torch.cuda.current_stream(self.device).cuda_stream
        self.x[f].zero_()
        torch.mul(self.b[f], 1.0, out=self.r[f])
        check(lib().hasty_cg_init(self.r[f].data_ptr(), self.p[f].data_ptr(), self.nvox, 0.0,
                                  self.state[f].data_ptr(), s))

    def _iteration(self, f: int) -> None:
        s = torch.cuda.current_stream(self.device).cuda_stream
        self.ops[f].apply(self.p[f], self.Ap[f])
        check(lib().hasty_cg_step(self.x[f].data_ptr(), self.r[f].data_ptr(), self.p[f].data_ptr(),
                                  self.Ap[f].data_ptr(), self.nvox, self.state[f].data_ptr(), s))

    def _rhs_branches(self) -> None:
        """Every frame's right-hand side, frame f on branch f mod streams."""
        main = torch.cuda.current_stream(self.device)
        for i, st in enumerate(self.streams):
            st.wait_stream(main)
            with torch.cuda.stream(st):
                for f in range(i, self.F, self.nstreams):
                    self._rhs(f)
        for st in self.streams:
            main.wait_stream(st)

    def _solve_batched(self, iters: int, rhs: bool) -> None:
        """CG of all frames as one batched system: per iteration one frames
        apply and one batched CG update."""
        if rhs:
            self._rhs_branches()
        s = torch.cuda.current_stream(self.device).cuda_stream
        self.x.zero_()
        torch.mul(self.b, 1.0, out=self.r)
        check(lib().hasty_cg_init_batched(self.r.data_ptr(), self.p.data_ptr(), self.nvox, self.F, 0.0,
                                          self.state.data_ptr(), s))
        for _ in range(iters):
            self.frames_op.apply(self.p, self.Ap)
            check(lib().hasty_cg_step_batched(self.x.data_ptr(), self.r.data_ptr(), self.p.data_ptr(),
                                              self.Ap.data_ptr(), self.nvox, self.F, self.state.data_ptr(), s))

    def _branch(self, i: int, iters: int, rhs: bool) -> None:
        """Frames i, i + streams, ...: right-hand side, CG init, iterations."""
        for f in range(i, self.F, self.nstreams):
            if rhs:
                self._rhs(f)
            self._init(f)
            for _ in range(iters):
                self._iteration(f)

    def _solve_launches(self, iters: int, branches: bool, rhs: bool) -> None:
        if not branches:
            for f in range(self.F):
                if rhs:
                    self._rhs(f)
                self._init(f)
                for _ in range(iters):
                    self._iteration(f)
            return
        main = torch.cuda.current_stream(self.device)
        for i, st in enumerate(self.streams):
            st.wait_stream(main)
            with torch.cuda.stream(st):
                self._branch(i, iters, rhs)
        for st in self.streams:
            main.wait_stream(st)

    def solve(self, iters: int, graph: bool = True, branches: bool = True, rhs: bool = True,
              batched: bool = True) -> torch.Tensor:
        """``iters`` CG iterations of every frame from x = 0; returns x [F, N...].

        ``rhs``: recompute the right-hand sides from the k-space data first
        (part of the solve; False reuses those of the last solve).
        ``batched``: the CG of all frames as one batched system (one launch
        per pass over all frames); otherwise per-frame launches.
        ``graph``: replay the captured solve (captured on first use for these
        arguments; the capture itself executes nothing).  ``branches``: run
        the per-frame launches on the concurrent streams (eager per-frame
        path; a per-frame graph always does).
        """
        if iters < 0:
            raise RuntimeError("iters must be >= 0")
        if not graph:
            if batched:
                self._solve_batched(iters, rhs)
            else:
                self._solve_launches(iters, branches, rhs)
            return self.x
        key = (int(iters), bool(rhs), bool(batched))
        g = self.graphs.get(key)
        if g is None:
            torch.cuda.synchronize(self.device)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                if batched:
                    self._solve_batched(iters, rhs)
                else:
                    self._solve_launches(iters, True, rhs)
            self.graphs[key] = g
        g.replay()
        return self.x

    def residuals(self) -> torch.Tensor:
        """sqrt(r^H r) per frame after the last iteration (host read)."""
        return self.state[:, 0].sqrt().cpu()
