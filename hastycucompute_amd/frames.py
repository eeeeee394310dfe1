"""Batched dynamic reconstruction on one GPU: frames with their own
(uneven) trajectories, reconstructed together.

Frame f has trajectory k_f (M_f points), its own Toeplitz PSF kernel K_f and
SENSE normal operator N_f x = (1/N^d) sum_c conj(S_c) A_f^H A_f (S_c x)
(``NORMAL_T_T1_OP``, all coils per launch, coil maps shared by the frames).
The right-hand sides b_f = (1/N^d) sum_c conj(S_c) A_f^H y_f come from one
all-coil type-1 NUFFT per frame.

CG runs on every frame at once.  One iteration of all frames -- F normal
operator applies and F fused CG updates, every scalar on the device -- is
captured once into a HIP graph (``torch.cuda.CUDAGraph``) and replayed, so
the host issues one graph launch per iteration and never synchronises
between iterations.  The eager path runs the same launches in the same
order; both give identical results (every kernel is deterministic).
"""
from __future__ import annotations

from typing import Dict, List, Optional, Sequence

import torch

from . import fft as hf
from ._lib import CG_STATE_DOUBLES, check, lib
from .mri import NORMAL_T_T1_OP


class FrameBatch:
    def __init__(self, smaps: torch.Tensor, coords: Sequence[Sequence[torch.Tensor]],
                 kdata: Sequence[torch.Tensor], tol: float = 1e-5, kernel_tol: float = 1e-9,
                 nufft_opts: Optional[Dict] = None):
        if smaps.dtype != torch.complex64 or smaps.device.type != "cuda":
            raise RuntimeError("smaps must be a complex64 GPU tensor [C, N...]")
        if len(coords) != len(kdata) or not coords:
            raise RuntimeError("one trajectory and one k-space array per frame")
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
        adj = hf.nufft_plan.make(hf.nufft_opts(nmodes=self.n, ntransf=self.C, tol=tol,
                                               sign=hf.nufft_sign.DEFAULT_TYPE_1, **opts),
                                 hf.nufft_type.TYPE_1)
        self.kernels: List[torch.Tensor] = []
        self.ops: List[NORMAL_T_T1_OP] = []
        self.b = torch.empty((self.F,) + self.n, dtype=torch.complex64, device=dev)
        img = torch.empty((self.C,) + self.n, dtype=torch.complex64, device=dev)
        for f in range(self.F):
            kf = [c.contiguous() for c in coords[f]]
            kernel = torch.empty(tuple(2 * s for s in self.n), dtype=torch.complex64, device=dev)
            hf.toeplitz_kernel(kf, kernel, tol=kernel_tol)
            self.kernels.append(kernel)
            self.ops.append(NORMAL_T_T1_OP(kernel, self.smaps))
            adj.setpts(kf)
            y = kdata[f].contiguous()
            if tuple(y.shape) != (self.C, int(kf[0].shape[0])):
                raise RuntimeError(f"kdata[{f}] must have shape [{self.C}, {int(kf[0].shape[0])}]")
            adj.execute(y, img)
            torch.sum(self.smaps.conj() * img, dim=0, out=self.b[f])
            self.b[f] /= float(self.nvox)
        adj.free()
        # CG workspace of all frames
        shape = (self.F,) + self.n
        self.x = torch.zeros(shape, dtype=torch.complex64, device=dev)
        self.r = torch.empty_like(self.x)
        self.p = torch.empty_like(self.x)
        self.Ap = torch.empty_like(self.x)
        self.state = torch.zeros((self.F, CG_STATE_DOUBLES), dtype=torch.float64, device=dev)
        self.graph: Optional[torch.cuda.CUDAGraph] = None

    def _stream(self):
        return torch.cuda.current_stream(self.device).cuda_stream

    def _init(self):
        self.x.zero_()
        self.r.copy_(self.b)
        s = self._stream()
        n = self.nvox
        for f in range(self.F):
            check(lib().hasty_cg_init(self.r[f].data_ptr(), self.p[f].data_ptr(), n, 0.0,
                                      self.state[f].data_ptr(), s))

    def _iteration(self):
        """One CG iteration of every frame (kernel launches only: capturable)."""
        s = self._stream()
        n = self.nvox
        for f in range(self.F):
            self.ops[f].apply(self.p[f], self.Ap[f])
            check(lib().hasty_cg_step(self.x[f].data_ptr(), self.r[f].data_ptr(), self.p[f].data_ptr(),
                                      self.Ap[f].data_ptr(), n, self.state[f].data_ptr(), s))

    def solve(self, iters: int, graph: bool = True) -> torch.Tensor:
        """``iters`` CG iterations of every frame from x = 0; returns x [F, N...]."""
        self._init()
        if iters <= 0:
            return self.x
        if not graph:
            for _ in range(iters):
                self._iteration()
            return self.x
        # the first iteration runs eagerly (it also warms every plan-owned
        # cache), the rest replay one captured iteration
        self._iteration()
        if self.graph is None and iters > 1:
            torch.cuda.synchronize(self.device)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._iteration()
            self.graph = g
            # capture records without executing: the state is as after one iteration
        for _ in range(iters - 1):
            self.graph.replay()
        return self.x

    def residuals(self) -> torch.Tensor:
        """sqrt(r^H r) per frame after the last iteration (host read)."""
        return self.state[:, 0].sqrt().cpu()
