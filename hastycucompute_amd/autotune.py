"""Autotuner for the NUFFT plan's tiling knobs on the GPU it runs on.

For a given transform (type, modes, transforms, tolerance) and trajectory it
times plans built with each candidate option set -- fine-grid tile per axis
(``tile``), points per spreading / interpolation block (``max_subproblem``)
and transforms per block (``coil_group``) -- and returns the fastest.  Every
candidate computes the same transform (the knobs only change how the work is
blocked), so the choice never changes results beyond rounding order.

Results are cached per (device, type, modes, transforms, tolerance, point
count rounded to a power of two) in memory and, when ``HASTY_AUTOTUNE_CACHE``
names a file, in that JSON file across processes.
"""
from __future__ import annotations

import json
import math
import os
from typing import Dict, List, Optional, Sequence

import torch

from . import fft as hf

_MEM: Dict[str, dict] = {}


def _candidates(nt: int, dim: int) -> List[dict]:
    if nt == 1:
        tiles = [(8, 8, 8), (8, 8, 16), (4, 8, 8), (8, 16, 16)] if dim == 3 else [(1, 16, 16), (1, 16, 32), (1, 32, 32)]
        subs = [1024, 2048, 4096]
        return [dict(tile=t, max_subproblem=s) for t in tiles for s in subs]
    tiles = [(16, 16, 16), (8, 16, 16), (16, 16, 8)] if dim == 3 else [(1, 32, 32), (1, 16, 32), (1, 64, 64)]
    subs = [4096, 8192, 16384]
    return [dict(tile=t, max_subproblem=s) for t in tiles for s in subs]


def _key(nt, nmodes, ntransf, tol, M, device) -> str:
    bucket = 1 << max(0, int(math.ceil(math.log2(max(M, 1)))))
    name = torch.cuda.get_device_name(device)
    return f"{name}|t{int(nt)}|{'x'.join(map(str, nmodes))}|b{ntransf}|tol{tol:g}|M{bucket}"


def _load_disk() -> Dict[str, dict]:
    path = os.environ.get("HASTY_AUTOTUNE_CACHE")
    if not path or not os.path.exists(path):
        return {}
    try:
        with open(path) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def _store_disk(key: str, value: dict) -> None:
    path = os.environ.get("HASTY_AUTOTUNE_CACHE")
    if not path:
        return
    data = _load_disk()
    data[key] = value
    tmp = path + ".tmp"
    with open(tmp, "w") as fh:
        json.dump(data, fh, indent=1)
    os.replace(tmp, path)


def _time_plan(nmodes, ntransf, tol, nt, opts, coords, data_in, data_out, reps) -> float:
    sign = hf.nufft_sign.DEFAULT_TYPE_1 if nt == hf.nufft_type.TYPE_1 else hf.nufft_sign.DEFAULT_TYPE_2
    plan = hf.nufft_plan.make(hf.nufft_opts(nmodes=tuple(nmodes), ntransf=ntransf, tol=tol, sign=sign,
                                            **opts), nt)
    try:
        plan.setpts(coords)
        plan.execute(data_in, data_out)            # warm (first-use allocations)
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        start.record()
        for _ in range(reps):
            plan.execute(data_in, data_out)
        stop.record()
        stop.synchronize()
        return start.elapsed_time(stop) / reps
    finally:
        plan.free()


def tune_nufft(nmodes: Sequence[int], ntransf: int, coords: Sequence[torch.Tensor], nt,
               tol: float = 1e-5, candidates: Optional[List[dict]] = None, reps: int = 3,
               use_cache: bool = True) -> dict:
    """Fastest option set (keyword arguments of ``nufft_opts``) for this
    transform and trajectory, with its time in ``best['ms']`` and every
    candidate's time in ``best['trials']``."""
    nt = hf.nufft_type(nt)
    dev = coords[0].device
    M = int(coords[0].shape[0])
    key = _key(nt, nmodes, ntransf, tol, M, dev)
    if use_cache:
        if key in _MEM:
            return dict(_MEM[key])
        disk = _load_disk()
        if key in disk:
            _MEM[key] = disk[key]
            return dict(disk[key])
    cands = candidates if candidates is not None else _candidates(int(nt), len(nmodes))
    img = torch.zeros((ntransf,) + tuple(nmodes), dtype=torch.complex64, device=dev)
    nu = torch.zeros((ntransf, M), dtype=torch.complex64, device=dev)
    g = torch.Generator(device=dev).manual_seed(5)
    if nt == hf.nufft_type.TYPE_1:
        nu.real.normal_(generator=g)
        data_in, data_out = nu, img
    else:
        img.real.normal_(generator=g)
        data_in, data_out = img, nu
    trials = []
    for c in cands:
        opts = {k: (list(v) if isinstance(v, tuple) else v) for k, v in c.items()}
        ms = _time_plan(nmodes, ntransf, tol, nt, opts, coords, data_in, data_out, reps)
        trials.append({"opts": opts, "ms": ms})
    best = min(trials, key=lambda t: t["ms"])
    result = dict(best["opts"])
    result["ms"] = best["ms"]
    result["trials"] = trials
    _MEM[key] = result
    _store_disk(key, result)
    return dict(result)


def plan_options(tuned: dict) -> dict:
    """The ``nufft_opts`` keyword arguments of a tuning result."""
    return {k: (tuple(v) if isinstance(v, list) else v) for k, v in tuned.items()
            if k in ("tile", "max_subproblem", "coil_group", "max_batch", "method")}
