#!/usr/bin/env python
"""Benchmark: A^H A applies/s + NUFFT Mpts/s on one MI355X.

Workload (default): 128^3 complex64 image, 32 coils, 3-D radial kooshball of
32768 spokes x 256 samples (8 388 608 points per coil), Toeplitz grid 256^3,
NUFFT eps 1e-5 (KB width 6, sigma 2).  Inputs are synthetic and deterministic
(hastycucompute_amd.synth) and resident in HBM before any timed region.

Timed phases (each: `warmup` untimed iterations, then exactly `steps`
iterations bracketed by torch.cuda.synchronize()):
  * aha   -- one SENSE Toeplitz normal-operator apply over all coils (value)
  * type2 -- one forward NUFFT of all coils (Mpts/s = C*M / t)
  * type1 -- one adjoint NUFFT of all coils
  * plain -- the same three on the plain per-coil path (one coil per launch
             sequence / one single-transform plan looped over the coils)
  * frames -- batched dynamic frames: 20 frames of 96^3 x 16 coils with
             uneven trajectories, 15 CG iterations each, eager vs hipGraph
Per-kernel launch times come from HIP events the library records on its own
stream (hasty_profile_*), used for the roofline of the dominant kernel.
Accuracy is checked in the same run in FP64 on the GPU (direct NDFT at a
subset of points) and reported per result.

Prints ONE JSON line.  --gpus N>1 is refused: this build runs on one GPU.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (N, coils, spokes, samples)
    "sense3d_128c32_koosh8M": (128, 32, 32768, 256),
    "sense3d_64c8_koosh131k": (64, 8, 1024, 128),
}
DEFAULT = "sense3d_128c32_koosh8M"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(N, C, M):
    """Algorithmic HBM bytes (c64 = 8 B) per launch of each hot kernel and per
    whole operation.  I = N^3 image voxels, G = (2N)^3 grid cells."""
    I = N ** 3
    G = 8 * I
    b = 8
    return {
        # Toeplitz passes (all coils per launch)
        "toep_P1_x_pad_fft": b * (I + C * I + 2 * C * I),      # x, S in; [C,N,N,2N] out
        "toep_P2_y_pad_fft": b * (2 * C * I + 4 * C * I),      # [C,N,N,2N] in; [C,N,2N,2N] out
        "toep_P3_z_psf": b * (4 * C * I + G + 4 * C * I),      # columns in, kernel, columns out
        "toep_P4_y_ifft_crop": b * (4 * C * I + 2 * C * I),
        "toep_P5_x_ifft_crop": b * (2 * C * I + C * I + I),    # rows, S in; image out
        # fused last pass (replaces P4 + P5 on 128 x 128 planes): [C,N,2N,2N] and S in, image out
        "toep_F45_yx_ifft_crop_sum": b * (4 * C * I + C * I + I),
        # fused first pass (replaces P1 + P2): x and S in, [C,N,2N,2N] out
        "toep_F12_xy_pad_fft": b * (I + C * I + 4 * C * I),
        # NUFFT kernels
        "nufft_interp": b * (C * G + C * M) + 4 * 3 * M + 4 * M * 7,   # grid, out, coords, sorted pts
        "nufft_spread": b * (C * G + C * M) + 4 * 3 * M + 4 * M * 7,
        # whole operations (model of the path's minimum traffic)
        "aha": b * (2 * I + 2 * C * I + G + 2 * C * G),
        "nufft": b * (I + C * I + 2 * C * G + C * M) + 4 * 3 * M,
    }


def fp64_ndft_type2_subset(img, coords, idx):
    """c_j = sum_n img[n] exp(-i k_j . n) in FP64 on the GPU for points idx."""
    import torch
    N = img.shape
    n = [torch.arange(-(s // 2), (s + 1) // 2, device=img.device, dtype=torch.float64) for s in N]
    k = [c[idx].to(torch.float64) for c in coords]
    e = [torch.exp(-1j * torch.outer(k[a], n[a])) for a in range(3)]
    f = img.to(torch.complex128)
    out = []
    for j0 in range(0, len(idx), 64):
        j1 = min(len(idx), j0 + 64)
        t = torch.einsum("zyx,jx->jzy", f, e[2][j0:j1].contiguous())
        t = torch.einsum("jzy,jy->jz", t, e[1][j0:j1].contiguous())
        out.append(torch.einsum("jz,jz->j", t, e[0][j0:j1].contiguous()))
    return torch.cat(out)


def fp64_direct_type1_subset(c, coords, shape, vox):
    """f[n] = sum_j c_j exp(+i k_j . n) in FP64 on the GPU for the centred mode
    indices n = vox [V, 3] (all M points, in chunks of 2^20)."""
    import torch
    n = vox.to(torch.float64)
    out = torch.zeros(vox.shape[0], dtype=torch.complex128, device=c.device)
    M = c.shape[0]
    for j0 in range(0, M, 1 << 20):
        j1 = min(M, j0 + (1 << 20))
        ph = sum(coords[a][j0:j1].to(torch.float64)[:, None] * n[None, :, a] for a in range(3))
        out += c[j0:j1].to(torch.complex128) @ torch.polar(torch.ones_like(ph), ph)
    return out


def fp64_plan_pair_normal(coords, smaps, x, N, tol=1e-9):
    """(1/N^3) sum_c conj(S_c) A1(A2(S_c x)) with complex128 NUFFT plans at tol."""
    import torch
    from hastycucompute_amd import fft as hf
    C = smaps.shape[0]
    c64 = [c.to(torch.float64) for c in coords]
    img = (smaps.to(torch.complex128) * x.to(torch.complex128)[None]).contiguous()
    opts = dict(nmodes=(N, N, N), ntransf=C, tol=tol)
    q2 = hf.nufft_plan.make(hf.nufft_opts(sign=-1, **opts), hf.nufft_type.TYPE_2, torch.float64)
    q2.setpts(c64)
    y = torch.empty((C, c64[0].shape[0]), dtype=torch.complex128, device=x.device)
    q2.execute(img, y)
    q2.free()
    q1 = hf.nufft_plan.make(hf.nufft_opts(sign=+1, **opts), hf.nufft_type.TYPE_1, torch.float64)
    q1.setpts(c64)
    q1.execute(y, img)
    q1.free()
    del y
    return (torch.conj(smaps.to(torch.complex128)) * img).sum(0) / float(N ** 3)


def cpu_baseline_nufft(N, C, coords_np, smaps_np, x_np, sample_coils):
    """Time the oracle's CPU restatement of the NUFFT (the KB gridding step in
    C with OpenMP, scipy.fft) on every host core, for a sample of the coils
    over the whole trajectory; Mpts/s extrapolated to all coils."""
    import numpy as np

    from oracle.kb import KBPlan
    cores = len(os.sched_getaffinity(0))
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    plan = KBPlan((N, N, N), 1e-5, dtype=np.complex64, workers=cores).setpts(coords_np)
    f = (smaps_np[:sample_coils] * x_np[None]).astype(np.complex64)
    t0 = time.perf_counter()
    yk = plan.type2(f)
    t2 = time.perf_counter() - t0
    t0 = time.perf_counter()
    plan.type1(yk)
    t1 = time.perf_counter() - t0
    M = len(coords_np[0])
    scale = C / sample_coils
    return {"type2_mpts_s": round(C * M / (t2 * scale) / 1e6, 2),
            "type1_mpts_s": round(C * M / (t1 * scale) / 1e6, 2),
            "unit": "Mpts/s", "cores": cores, "kind": "port",
            "sample": f"oracle KB NUFFT (complex64, C gridding with OpenMP + scipy.fft, {cores} threads), "
                      f"{sample_coils} of {C} coils over all {M} points at {N}^3, one run each, "
                      f"extrapolated x{scale:g}",
            "seconds": {"type2": round(t2, 3), "type1": round(t1, 3)}}


def cpu_baseline(N, C, smaps_np, x_np, kernel_np, sample_coils, reps):
    """Time the oracle's CPU restatement of the Toeplitz normal operator on a
    bounded sample (``reps`` applies over ``sample_coils`` coils; complex64,
    scipy.fft on every host core)."""
    import numpy as np

    from oracle.solvers import sense_normal
    cores = len(os.sched_getaffinity(0))
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    S = smaps_np[:sample_coils]
    sense_normal(x_np, S[:1], kernel_np, dtype=np.complex64, workers=cores)   # warm
    t0 = time.perf_counter()
    for _ in range(reps):
        sense_normal(x_np, S, kernel_np, dtype=np.complex64, workers=cores)
    t = (time.perf_counter() - t0) / reps
    per_apply = t * C / sample_coils
    return {"value": 1.0 / per_apply, "unit": "applies/s", "cores": cores, "kind": "port",
            "sample": f"oracle Toeplitz normal op (complex64, scipy.fft, {cores} threads) on "
                      + (f"all {C} coils" if sample_coils == C else f"{sample_coils} of {C} coils")
                      + f" at {N}^3, {reps} applies"
                      + ("" if sample_coils == C else f", extrapolated x{C / sample_coils:g}"),
            "seconds_per_apply": per_apply}


# Kernels whose HBM traffic is measured with rocprofv3 --pmc, by the symbol
# substring that identifies their launches (the Toeplitz passes are the
# two-level FFT kernel in its column-PSF mode, unique to P3), plus the
# calibration kernel: torch.neg over 1 GiB reads and writes exactly 2^30 B.
PMC_KERNELS = {
    "toep_P3_z_psf": "fft2_pass_kernel<HIP_vector_type<float, 2u>, 16, 16, -1, 1, 3>",
    "toep_F45_yx_ifft_crop_sum": "toep_crop_sum_kernel",
    "toep_F12_xy_pad_fft": "toep_pad_fft_kernel",
    "nufft_spread": "spread_lanes_kernel<6, 32>",
    "nufft_interp": "interp_pair_kernel<6, 16, 16, 16>",
}
CALIB_SUBSTR = "neg_kernel"
CALIB_BYTES = 1 << 30


def _pmc_rows(d):
    """(kernel name, counter value) rows of every counter_collection CSV under d."""
    import csv
    import glob
    rows = []
    for f in glob.glob(os.path.join(d, "**", "*counter_collection*.csv"), recursive=True):
        with open(f) as fh:
            for r in csv.DictReader(fh):
                name = r.get("Kernel_Name") or r.get("Kernel-Name") or ""
                val = r.get("Counter_Value") or r.get("Counter-Value")
                if val is not None:
                    rows.append((name, float(val)))
    return rows


def reduce_pmc(fetch_dir, write_dir, workload):
    """Average FETCH_SIZE / WRITE_SIZE (KiB per dispatch) of each measured
    kernel, scaled by the factor that maps the calibration kernel's reading to
    its known 1 GiB, into bytes per launch."""
    cmd = ("cd /tmp && export TMPDIR=/tmp && cd $REPO && rocprofv3 --pmc {c} --output-format csv "
           "-d {d} -o run -- python -u bench.py --steps 1 --warmup 0 --no-cpu-baseline "
           "--phases aha,type2,type1")
    out = {"workload": workload,
           "commands": [cmd.format(c="FETCH_SIZE", d=fetch_dir), cmd.format(c="WRITE_SIZE", d=write_dir),
                        f"python bench.py --reduce-pmc {fetch_dir} {write_dir}"],
           "counters_kib_per_dispatch": {}, "calibration": {}, "kernels": {}}
    scaled = {}
    for counter, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        rows = _pmc_rows(d)
        cal = [v for n, v in rows if CALIB_SUBSTR in n]
        if not cal:
            raise SystemExit(f"no calibration dispatch ({CALIB_SUBSTR}) in {d}")
        factor = CALIB_BYTES / (max(cal) * 1024.0)
        out["calibration"][counter] = {"kib_measured": max(cal), "bytes_known": CALIB_BYTES,
                                       "factor": factor}
        for tag, sub in PMC_KERNELS.items():
            vals = [v for n, v in rows if sub in n]
            if vals:
                kib = sum(vals) / len(vals)
                out["counters_kib_per_dispatch"].setdefault(tag, {})[counter] = kib
                scaled.setdefault(tag, {})[counter] = kib * 1024.0 * factor
    for tag, c in scaled.items():
        if "FETCH_SIZE" in c and "WRITE_SIZE" in c:
            out["kernels"][tag] = {"fetch_bytes": c["FETCH_SIZE"], "write_bytes": c["WRITE_SIZE"],
                                   "traffic_bytes": c["FETCH_SIZE"] + c["WRITE_SIZE"]}
    path = os.path.join(ROOT, "profiles", f"pmc_traffic_{workload}.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out, indent=1))


_TRACE_CMD = ("cd /tmp && export TMPDIR=/tmp && cd $REPO && rocprofv3 --kernel-trace --stats --output-format csv "
              "-d {d} -o run -- python -u bench.py --steps 3 --warmup 1 --no-cpu-baseline")
# per-kernel HIP-event timing: launches of each operator, one operator after
# another (consecutive launches of one operator, as in the timed phases)
PROFILE_LAUNCHES = 10
# the launches this script times after its calibration kernel: the apply, then
# type 2, then type 1, PROFILE_LAUNCHES times each, as (label, kernel-name substring)
_TIMED_SEQ = ([("toep_F12_xy_pad_fft", "toep_pad_fft_kernel"),
               ("toep_P3_z_psf", "<c64, 16, 16, -1, 1, 3>"), ("toep_F45_yx_ifft_crop_sum", "toep_crop_sum_kernel"),
               ("toep_F45 sum_partials", "sum_partials")] * PROFILE_LAUNCHES +
              [("fwd_xy_pad_fft", "nufft_fwd_xy_kernel"),
               ("fwd_z_pad_fft", "<c64, 16, 16, -1, 0, 1>"), ("nufft_interp", "interp_pair_kernel<6,")] * PROFILE_LAUNCHES +
              [("nufft_sort_values", "sort_values_kernel<c64"), ("nufft_spread", "spread_lanes_kernel<6, 32>"),
               ("nufft_spread_reduce", "reduce_split_tiles"), ("inv_z_fft_crop", "<c64, 16, 16, 1, 0, 4>"),
               ("inv_xy_fft_crop", "nufft_inv_xy_kernel")] * PROFILE_LAUNCHES)


def _short_kernel(name):
    import re
    name = name.replace("HIP_vector_type<float, 2u>", "c64").replace("HIP_vector_type<double, 2u>", "c128")
    name = re.sub(r"\(hasty::.*$", "", name)
    name = re.sub(r"\(.*$", "", name)
    return name.replace("void ", "")[:100]


def summarize_trace(trace_dir, workload, top=30):
    """Markdown summary of a rocprofv3 --kernel-trace run of this script:
    launches grouped by kernel and launch shape, then the timed launches."""
    import collections
    import csv
    import glob
    files = glob.glob(os.path.join(trace_dir, "**", "*kernel_trace*.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no kernel_trace CSV under {trace_dir}")
    rows = [r for f in files for r in csv.DictReader(open(f))]
    groups = collections.defaultdict(list)
    total = 0.0
    for r in rows:
        d = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6
        total += d
        key = (_short_kernel(r["Kernel_Name"]), f'{r["Grid_Size_X"]}x{r["Grid_Size_Y"]}', r["Workgroup_Size_X"],
               r["VGPR_Count"], r["LDS_Block_Size"])
        groups[key].append(d)
    N, C, spokes, samples = WORKLOADS[workload]
    out = [f"# Kernel trace: {workload} ({N}^3 image, {C} coils, {spokes * samples / 1e6:.1f}M kooshball points "
           f"per coil, {2 * N}^3 Toeplitz grid)", "",
           "Command (from the repository root, on one MI355X), then the summary:", "", "```",
           _TRACE_CMD.format(d=trace_dir), f"python bench.py --summarize-trace {trace_dir}", "```", "",
           "Launches are grouped by kernel and launch shape (grid in work-items), so the full-size launches of a "
           "pass are separate from the same kernel's smaller launches (plain per-coil path, frames, checks). "
           "VGPR is the trace's VGPR_Count field.", "",
           "| kernel | grid | block | VGPR | LDS B | calls | avg ms | min ms | max ms | % of GPU time |",
           "|---|---|---:|---:|---:|---:|---:|---:|---:|---:|"]
    for (name, grid, wg, vgpr, lds), ds in sorted(groups.items(), key=lambda kv: -sum(kv[1]))[:top]:
        out.append(f"| `{name}` | {grid} | {wg} | {vgpr} | {lds} | {len(ds)} | {sum(ds) / len(ds):.4f} | "
                   f"{min(ds):.4f} | {max(ds):.4f} | {100 * sum(ds) / total:.2f} |")
    order = sorted(rows, key=lambda r: int(r["Start_Timestamp"]))
    mark = max(i for i, r in enumerate(order) if CALIB_SUBSTR in r["Kernel_Name"])
    timed = collections.OrderedDict()
    for (tag, pat), r in zip(_TIMED_SEQ, order[mark + 1:mark + 1 + len(_TIMED_SEQ)]):
        name = _short_kernel(r["Kernel_Name"])
        if pat not in name:
            raise SystemExit(f"unexpected launch {name} where {tag} was expected")
        timed.setdefault((tag, name), []).append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6)
    out += ["", "## Launches timed by bench.py's roofline", "",
            "The launches that follow the calibration kernel (`torch.neg` over 1 GiB): an A^H A apply (the fused x/y "
            f"pad-FFT pass, P3, the fused y/x inverse pass with the coil sum, and the coil-group partials), a type-2 and a type-1 transform at full size, {PROFILE_LAUNCHES} times each, one operator after another, which "
            "bench.py times with HIP events on the library stream. Several passes share a kernel instance, so "
            "each launch is labelled by its place in that sequence.", "",
            "| pass | kernel | launches | avg ms |", "|---|---|---:|---:|"]
    for (tag, name), ds in timed.items():
        out.append(f"| {tag} | `{name}` | {len(ds)} | {sum(ds) / len(ds):.4f} |")
    path = os.path.join(ROOT, "profiles", f"{workload}_kernel_stats.md")
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")
    print(path)


_SQ_PASSES = ("SQ_INSTS_VALU SQ_INSTS_SALU SQ_INSTS_SMEM SQ_INSTS_VMEM_RD SQ_WAIT_INST_ANY SQ_WAVE_CYCLES "
              "SQ_BUSY_CYCLES SQ_ACTIVE_INST_VALU",
              "SQ_WAVES SQ_INSTS_VALU_FMA_F32 SQ_INSTS_LDS SQ_LDS_BANK_CONFLICT SQ_ACTIVE_INST_LDS GRBM_GUI_ACTIVE",
              "SQ_INSTS_MFMA SQ_VALU_MFMA_BUSY_CYCLES SQ_INSTS_VMEM_WR SQ_WAIT_ANY")
_SQ_KERNELS = (("spread_lanes_kernel<6, 32>", "spread_lanes_kernel<6, 32>", 10_000_000),
               ("interp_pair_kernel<6, 16, 16, 16>", "interp_pair_kernel<6, 16, 16, 16>", 20_000_000))


# the Toeplitz passes of the A^H A apply (full-size launches: the grid of the
# fused passes is n_z x coil groups x threads, of P3 at least 2^20 work-items)
_SQ_TOEP_PASSES = ("SQ_INSTS_VALU SQ_INSTS_LDS SQ_WAIT_INST_ANY SQ_WAVE_CYCLES SQ_BUSY_CYCLES SQ_ACTIVE_INST_VALU "
                   "SQ_ACTIVE_INST_LDS SQ_LDS_BANK_CONFLICT",
                   "SQ_WAVES SQ_INSTS_VMEM_RD SQ_INSTS_VMEM_WR SQ_WAIT_ANY SQ_INSTS_SALU GRBM_GUI_ACTIVE")
_SQ_TOEP_KERNELS = (("toep_F12_xy_pad_fft", "toep_pad_fft_kernel", 100_000),
                    ("toep_P3_z_psf", "fft2_pass_kernel<HIP_vector_type<float, 2u>, 16, 16, -1, 1, 3>", 1_000_000),
                    ("toep_F45_yx_ifft_crop_sum", "toep_crop_sum_kernel", 100_000))


def summarize_sq(dirs, workload, toeplitz=False):
    """Markdown summary of the SQ-counter passes (one directory per pass)
    over the NUFFT spread and interpolation kernels, or (toeplitz) over the
    passes of the A^H A apply (full-size launches only)."""
    kernels = _SQ_TOEP_KERNELS if toeplitz else _SQ_KERNELS
    passes = _SQ_TOEP_PASSES if toeplitz else _SQ_PASSES
    if len(dirs) != len(passes):
        raise SystemExit(f"expected {len(passes)} pass directories, got {len(dirs)}")
    import collections
    import csv
    import glob
    res = collections.defaultdict(dict)
    for d in dirs:
        agg = collections.defaultdict(lambda: collections.defaultdict(list))
        for f in glob.glob(os.path.join(d, "**", "*counter_collection*.csv"), recursive=True):
            for r in csv.DictReader(open(f)):
                for label, sub, min_grid in kernels:
                    if sub in r["Kernel_Name"] and int(r["Grid_Size"]) >= min_grid:
                        agg[label][r["Counter_Name"]].append(float(r["Counter_Value"]))
        for label, counters in agg.items():
            for c, vals in counters.items():
                res[label][c] = sum(vals) / len(vals)
    cmd = ("cd /tmp && export TMPDIR=/tmp && cd $REPO && rocprofv3 --pmc {c} --output-format csv -d {d} -o run "
           "-- python -u bench.py --steps 1 --warmup 0 --no-cpu-baseline --phases "
           + ("aha" if toeplitz else "type2,type1"))
    labels = [k for k, _, _ in kernels if k in res]
    title = "Toeplitz passes of the A^H A apply" if toeplitz else "NUFFT spread and interpolation"
    flag = "--summarize-sq-toeplitz" if toeplitz else "--summarize-sq"
    out = [f"# SQ counters: {title}, {workload}", "",
           f"{len(passes)} counter passes (the hardware takes at most 8 SQ counters per pass), from the repository "
           "root on one MI355X, then the summary:", "", "```"] + [cmd.format(c=c, d=d) for c, d in zip(passes, dirs)] + [
           f"python bench.py {flag} {' '.join(dirs)}", "```", "",
           "Values are per launch (mean over the full-size launches), in units of 10^9. Instruction counts are "
           "wave-instructions summed over the GPU; `GRBM_GUI_ACTIVE` is summed over the 8 XCDs.", "",
           "| counter | " + " | ".join(f"`{k}`" for k in labels) + " |", "|---|" + "---:|" * len(labels)]
    for c in sorted({c for k in labels for c in res[k]}):
        out.append(f"| {c} | " + " | ".join(f"{res[k][c] / 1e9:.3f}" if c in res[k] else "" for k in labels) + " |")
    out.append("")
    for k in labels:
        v, f = res[k].get("SQ_INSTS_VALU"), res[k].get("SQ_INSTS_VALU_FMA_F32")
        mf, act = res[k].get("SQ_INSTS_MFMA"), res[k].get("GRBM_GUI_ACTIVE")
        if mf and act:
            # v_mfma_f32_32x32x2_f32: 32 x 32 x 2 multiply-adds per wave-instruction
            busy = res[k].get("SQ_VALU_MFMA_BUSY_CYCLES")
            line = (f"- `{k}`: {mf / 1e6:.0f} M MFMA wave-instructions = {mf * 4096 / 1e12:.2f} TFLOP issued "
                    f"per launch (zero weights included), over {act / 8 / 1e6:.2f} M GPU clock cycles "
                    f"(GRBM_GUI_ACTIVE per XCD)")
            if busy:
                # summed over the 1024 SIMDs (256 CUs x 4)
                line += f"; MFMA busy {100 * busy / (1024 * act / 8):.0f}% of the SIMD cycles"
            out.append(line + ".")
        if v and f:
            out.append(f"- `{k}`: VALU-issue floor {v / 256 / 2.4e9 * 1e3:.1f} ms (one wave-instruction per CU per "
                       f"cycle, 256 CUs, 2.4 GHz); FP32 FMAs are {100 * f / v:.0f}% of its VALU instructions.")
        elif v:
            out.append(f"- `{k}`: VALU-issue floor {v / 256 / 2.4e9 * 1e3:.2f} ms (one wave-instruction per CU per "
                       f"cycle, 256 CUs, 2.4 GHz).")
    name = "toeplitz_passes_sq_counters" if toeplitz else "nufft_spread_interp_sq_counters"
    path = os.path.join(ROOT, "profiles", f"{name}_{workload}.md")
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")
    print(path)


def measured_traffic(workload, tag):
    """HBM bytes per launch of kernel `tag` from the committed PMC summary, or None."""
    path = os.path.join(ROOT, "profiles", f"pmc_traffic_{workload}.json")
    try:
        with open(path) as fh:
            d = json.load(fh)
        return d["kernels"][tag]["traffic_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def bench_frames(dev, iters=15, nframes=20, N=96, C=16, samples=192, seed=3, reps=3):
    """Batched dynamic frames: ``nframes`` rotated golden-means kooshballs of
    U[500, 3000] spokes x ``samples`` at N^3 with C coils, 15 CG iterations
    per frame on each frame's Toeplitz normal operator.  Timed: eager
    per-frame launches (one stream, then concurrent branches), the per-frame
    solve replayed from one hipGraph, and the batched solve (one launch per
    pass over all frames), eager and from a graph.  Returns the timings."""
    import numpy as np
    import torch

    from hastycucompute_amd import autotune, synth
    from hastycucompute_amd import fft as hf
    from hastycucompute_amd.frames import FrameBatch

    rng = np.random.default_rng(seed)
    spokes = rng.integers(500, 3001, nframes)
    n = (N, N, N)
    S = torch.from_numpy(synth.coil_maps(n, C, seed=1)).to(dev)
    x = torch.from_numpy(synth.image(n, seed=0)).to(dev)
    coords = [[torch.from_numpy(c).to(dev) for c in synth.rotated_kooshball(int(s), samples, f)]
              for f, s in enumerate(spokes)]
    t0 = time.perf_counter()
    tuned = autotune.tune_nufft(n, C, coords[0], hf.nufft_type.TYPE_1, tol=1e-5, reps=2)
    tune_s = time.perf_counter() - t0
    fwd = hf.nufft_plan.make(hf.nufft_opts(nmodes=n, ntransf=C, tol=1e-5, sign=-1), hf.nufft_type.TYPE_2)
    kdata = []
    img = (S * x[None]).contiguous()
    for f in range(nframes):
        fwd.setpts(coords[f])
        y = torch.empty((C, int(coords[f][0].shape[0])), dtype=torch.complex64, device=dev)
        fwd.execute(img, y)
        kdata.append(y + 0.01 * y.abs().mean() * torch.randn_like(y))
    fwd.free()
    t0 = time.perf_counter()
    fb = FrameBatch(S, coords, kdata, tol=1e-5, nufft_opts=autotune.plan_options(tuned))
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    # one solve = every frame's right-hand side (type-1 NUFFT of its data) and
    # `iters` CG iterations; eager one frame after another on one stream, eager
    # on the concurrent branches, and the whole solve replayed from one graph
    modes = {"eager": dict(graph=False, branches=False, batched=False),
             "eager_branches": dict(graph=False, branches=True, batched=False),
             "graph": dict(graph=True, batched=False),
             "batched_eager": dict(graph=False, batched=True),
             "batched_graph": dict(graph=True, batched=True)}
    out = {}
    for mode, kw in modes.items():
        fb.solve(iters, **kw)       # warm (and capture)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            fb.solve(iters, **kw)
        torch.cuda.synchronize()
        out[mode] = (time.perf_counter() - t) / reps
    xe = fb.solve(iters, graph=False, branches=False, batched=False).clone()
    xg = fb.solve(iters, graph=True, batched=False).clone()
    xb = fb.solve(iters, graph=True, batched=True).clone()
    same = bool(torch.equal(xe, xg)) and bool(torch.equal(xe, xb))
    best = min(out["graph"], out["batched_graph"])
    total_pts = int(sum(int(c[0].shape[0]) for c in coords))
    log(f"[bench] frames: {nframes} x {N}^3 x {C} coils, {total_pts} pts, {iters} CG iters: "
        f"eager {out['eager'] * 1e3:.1f} ms, eager branches {out['eager_branches'] * 1e3:.1f} ms, "
        f"graph {out['graph'] * 1e3:.1f} ms, batched eager {out['batched_eager'] * 1e3:.1f} ms, "
        f"batched graph {out['batched_graph'] * 1e3:.1f} ms, identical={same}")
    return {"frames": nframes, "image": [N] * 3, "coils": C, "cg_iters": iters,
            "spokes_per_frame": [int(s) for s in spokes], "samples_per_spoke": samples,
            "total_points_per_coil": total_pts,
            "solve": "right-hand sides (type-1 NUFFT per frame) + CG, every frame",
            "streams": fb.nstreams,
            "eager_ms_per_solve": round(out["eager"] * 1e3, 3),
            "eager_branches_ms_per_solve": round(out["eager_branches"] * 1e3, 3),
            "graph_ms_per_solve": round(out["graph"] * 1e3, 3),
            "batched_eager_ms_per_solve": round(out["batched_eager"] * 1e3, 3),
            "batched_graph_ms_per_solve": round(out["batched_graph"] * 1e3, 3),
            "graph_speedup": round(out["eager"] / out["graph"], 3),
            "batched_graph_speedup": round(out["eager"] / out["batched_graph"], 3),
            "frame_iterations_per_s": round(nframes * iters / best, 1),
            "graph_equals_eager": same,
            "setup_s": round(setup_s, 2), "autotune_s": round(tune_s, 2),
            "autotuned_type1_opts": autotune.plan_options(tuned)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # (the apply reaches its steady rate after a few dozen launches: 2.28 ms
    # per apply over 10 steps after 3 warm-up steps, 2.18 over 50 after 10,
    # 2.15 over 100 after 20)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default=DEFAULT, choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    # CPU baselines: about 10 s of host work each (a full 32-coil apply takes
    # ~1 s on 16 cores; a type 2 + type 1 of 8 coils ~14 s)
    ap.add_argument("--cpu-sample-coils", type=int, default=32)
    ap.add_argument("--cpu-reps", type=int, default=10)
    ap.add_argument("--cpu-nufft-coils", type=int, default=8)
    ap.add_argument("--phases", default="aha,type2,type1,plain,segments,frames")
    ap.add_argument("--reduce-pmc", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="reduce two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) of this script "
                         "into the traffic summary profiles/pmc_traffic_<workload>.json and exit")
    ap.add_argument("--summarize-trace", metavar="TRACE_DIR",
                    help="summarise a rocprofv3 --kernel-trace run of this script into "
                         "profiles/<workload>_kernel_stats.md and exit")
    ap.add_argument("--summarize-sq", nargs=3, metavar=("PASS1_DIR", "PASS2_DIR", "PASS3_DIR"),
                    help="summarise the three SQ-counter passes over the NUFFT kernels into "
                         "profiles/nufft_spread_interp_sq_counters_<workload>.md and exit")
    ap.add_argument("--summarize-sq-toeplitz", nargs=2, metavar=("PASS1_DIR", "PASS2_DIR"),
                    help="summarise the two SQ-counter passes over the Toeplitz passes into "
                         "profiles/toeplitz_passes_sq_counters_<workload>.md and exit")
    args = ap.parse_args()
    if args.reduce_pmc:
        reduce_pmc(args.reduce_pmc[0], args.reduce_pmc[1], args.workload)
        return
    if args.summarize_trace:
        summarize_trace(args.summarize_trace, args.workload)
        return
    if args.summarize_sq:
        summarize_sq(args.summarize_sq, args.workload)
        return
    if args.summarize_sq_toeplitz:
        summarize_sq(args.summarize_sq_toeplitz, args.workload, toeplitz=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the coil-sharded multi-GPU design is documented, not built)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from hastycucompute_amd import _lib, synth
    from hastycucompute_amd import fft as hf
    from hastycucompute_amd.mri import NORMAL_T_T1_OP

    N, C, spokes, samples = WORKLOADS[args.workload]
    M = spokes * samples
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    phases = set(args.phases.split(","))

    log(f"[bench] workload {args.workload}: N={N} C={C} M={M}")
    t0 = time.perf_counter()
    coords_np = synth.kooshball(spokes, samples)
    smaps_np = synth.coil_maps((N, N, N), C, seed=1)
    x_np = synth.image((N, N, N), seed=0)
    coords = [torch.from_numpy(c).to(dev) for c in coords_np]
    smaps = torch.from_numpy(smaps_np).to(dev)
    x = torch.from_numpy(x_np).to(dev)
    log(f"[bench] inputs ready in {time.perf_counter() - t0:.1f}s")

    # PSF precompute (FP64 NUFFT + FP64 FFT), one-off
    kernel = torch.empty((2 * N,) * 3, dtype=torch.complex64, device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hf.toeplitz_kernel(coords, kernel)
    torch.cuda.synchronize()
    psf_s = time.perf_counter() - t0
    log(f"[bench] toeplitz_kernel (FP64) {psf_s:.2f}s")

    op = NORMAL_T_T1_OP(kernel, smaps)
    out = torch.empty_like(x)
    opts = dict(nmodes=(N, N, N), ntransf=C, tol=1e-5)
    p2 = hf.nufft_plan.make(hf.nufft_opts(sign=-1, **opts), hf.nufft_type.TYPE_2)
    p1 = hf.nufft_plan.make(hf.nufft_opts(sign=+1, **opts), hf.nufft_type.TYPE_1)
    t0 = time.perf_counter()
    p2.setpts(coords)
    p1.setpts(coords)
    torch.cuda.synchronize()
    setpts_s = (time.perf_counter() - t0) / 2
    img_c = (smaps * x[None]).contiguous()
    y = torch.empty((C, M), dtype=torch.complex64, device=dev)
    back = torch.empty_like(img_c)

    def run_phase(name, fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            fn()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / args.steps
        log(f"[bench] {name}: {dt * 1e3:.3f} ms/iter")
        return dt

    res = {}
    if "aha" in phases:
        res["aha"] = run_phase("aha", lambda: op.apply(x, out))

    # 30 CG iterations on the normal equations N x = b, b = N x_true (FP64
    # dot products and scalars on the device), timed right after the applies
    # they are made of, before the NUFFT phases
    cg30 = None
    if "aha" in phases:
        from hastycucompute_amd.op import conjugate_gradient
        b = torch.empty_like(x)
        op.apply(x, b)
        xs = torch.zeros_like(x)
        hist = []
        # warm with the residual history too: its first torch sqrt loads a
        # code object, which inside the timed run added 15-70 ms
        conjugate_gradient(op, 30, 0.0)(xs, b, history=[])
        xs.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        conjugate_gradient(op, 30, 0.0)(xs, b, history=hist)
        torch.cuda.synchronize()
        cg_s = time.perf_counter() - t0
        r_true = torch.empty_like(x)
        op.apply(xs, r_true)
        bn = float(b.to(torch.complex128).norm())
        cg30 = {"iters": 30, "ms": round(cg_s * 1e3, 2), "ms_per_iter": round(cg_s * 1e3 / 30, 3),
                "recurrence_residual_rel": float(hist[-1].item()) / bn,
                "true_residual_rel": float((b - r_true).to(torch.complex128).norm()) / bn,
                "error_vs_truth_rel": float((xs - x).to(torch.complex128).norm() / x.to(torch.complex128).norm())}
        log(f"[bench] CG 30 iterations: {cg30}")

    if "type2" in phases:
        res["type2"] = run_phase("type2", lambda: p2.execute(img_c, y))
    if "type1" in phases:
        res["type1"] = run_phase("type1", lambda: p1.execute(y, back))

    # the plain per-coil GPU path, as the reference runs it: one coil per
    # normal-operator launch sequence, one single-transform NUFFT plan looped
    # over the coils (nufft.cppm:316-382 style, without its host copies)
    plain = {}
    if "plain" in phases:
        op1 = NORMAL_T_T1_OP(kernel, smaps, coil_batch=1)
        plain["aha"] = run_phase("plain per-coil aha", lambda: op1.apply(x, out))
        del op1
        one = dict(nmodes=(N, N, N), ntransf=1, tol=1e-5)
        q2 = hf.nufft_plan.make(hf.nufft_opts(sign=-1, **one), hf.nufft_type.TYPE_2)
        q1 = hf.nufft_plan.make(hf.nufft_opts(sign=+1, **one), hf.nufft_type.TYPE_1)
        q2.setpts(coords)
        q1.setpts(coords)
        y_views = [y[c:c + 1] for c in range(C)]
        img_views = [img_c[c:c + 1] for c in range(C)]
        back_views = [back[c:c + 1] for c in range(C)]

        def per_coil(plan, ins, outs):
            for a, b in zip(ins, outs):
                plan.execute(a, b)
        plain["type2"] = run_phase("plain per-coil type2", lambda: per_coil(q2, img_views, y_views))
        plain["type1"] = run_phase("plain per-coil type1", lambda: per_coil(q1, y_views, back_views))
        q1.free()
        q2.free()

    # an elementwise kernel of known traffic (reads 1 GiB, writes 1 GiB), for
    # scaling rocprof's FETCH_SIZE / WRITE_SIZE readings
    calib = torch.zeros(CALIB_BYTES // 4, dtype=torch.float32, device=dev)
    calib_dst = torch.empty_like(calib)
    torch.neg(calib, out=calib_dst)
    torch.cuda.synchronize()
    del calib, calib_dst

    # per-kernel HIP-event timing (library stream), PROFILE_LAUNCHES
    # consecutive launches of each operator
    _lib.profile_enable(True)
    for _ in range(PROFILE_LAUNCHES):
        op.apply(x, out)
    for _ in range(PROFILE_LAUNCHES):
        p2.execute(img_c, y)
    for _ in range(PROFILE_LAUNCHES):
        p1.execute(y, back)
    kern = _lib.profile_read()
    _lib.profile_enable(False)
    kern_avg = {k: (ms / n, n) for k, (ms, n) in kern.items()}
    for k, (ms, n) in sorted(kern_avg.items(), key=lambda kv: -kv[1][0]):
        log(f"[bench]   kernel {k}: {ms:.3f} ms avg over {n}")

    # accuracy in FP64 on the GPU
    acc = {}
    g = torch.Generator().manual_seed(7)
    idx = torch.randperm(M, generator=g)[:256].to(dev)
    ref = fp64_ndft_type2_subset(img_c[0], coords, idx)
    got = y[0][idx].to(torch.complex128)
    acc["type2_vs_fp64_ndft_256pts"] = float((got - ref).norm() / ref.norm())
    # A^H A vs the NUFFT pair (1/N^3) sum_c conj(S_c) A1(A2(S_c x))
    p2.execute(img_c, y)
    p1.execute(y, back)
    pair = (torch.conj(smaps) * back).sum(0) / float(N ** 3)
    op.apply(x, out)
    acc["aha_vs_nufft_pair"] = float((out - pair).norm() / pair.norm())
    del pair
    # type 1 vs an FP64 direct sum on 256 random voxels, coils 0 and C - 1
    vox = torch.stack([torch.randint(-(N // 2), (N + 1) // 2, (256,), generator=g) for _ in range(3)], 1).to(dev)
    vidx = tuple((vox[:, a] + N // 2) for a in range(3))
    errs = []
    for cc in (0, C - 1):
        ref1 = fp64_direct_type1_subset(y[cc], coords, (N, N, N), vox)
        errs.append(float((back[cc][vidx].to(torch.complex128) - ref1).norm() / ref1.norm()))
    acc["type1_vs_fp64_direct_256vox"] = max(errs)
    # the whole A^H A vs complex128 NUFFT plans at tol 1e-9 (FP64 arithmetic)
    if "aha" in phases:
        ref_n = fp64_plan_pair_normal(coords, smaps, x, N)
        op.apply(x, out)
        acc["aha_vs_fp64_plans_tol1e-9"] = float((out.to(torch.complex128) - ref_n).norm() / ref_n.norm())
        del ref_n
        torch.cuda.empty_cache()
    log(f"[bench] accuracy: {acc}")



    ab = algorithmic_bytes(N, C, M)
    aha_t = res.get("aha")
    value = 1.0 / aha_t if aha_t else None
    # dominant kernel among the A^H A passes by measured time
    toep = {k: v for k, v in kern_avg.items() if k.startswith("toep_")}
    roof = None
    if toep:
        dom = max(toep, key=lambda k: toep[k][0])
        t_dom = toep[dom][0] * 1e-3
        achieved = ab[dom] / t_dom / 1e9
        traffic = measured_traffic(args.workload, dom)
        roof = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": round(traffic) if traffic else None,
                "algorithmic_bytes_per_launch": ab[dom], "avg_launch_ms": round(toep[dom][0], 4)}
    # the NUFFT half of the metric: HBM view of interpolation and spreading,
    # and the matrix-core view of spreading (4 C M w^3 useful flops: a
    # complex value times a real weight, added, per coil per window cell --
    # two FP32 multiply-adds -- run as FP32 MFMA)
    roof_nufft = {}
    w = p2.info()["width"]
    for tag, kname in (("type2", "nufft_interp"), ("type1", "nufft_spread")):
        if kname in kern_avg:
            t = kern_avg[kname][0] * 1e-3
            achieved = ab[kname] / t / 1e9
            traffic = measured_traffic(args.workload, kname)
            e = {"kernel": kname, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                 "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                 "traffic": round(traffic) if traffic else None, "algorithmic_bytes_per_launch": ab[kname],
                 "avg_launch_ms": round(kern_avg[kname][0], 4)}
            if kname == "nufft_spread":
                useful = 4.0 * C * M * w ** 3
                e["useful_tflops"] = round(useful / t / 1e12, 2)
                e["frac_of_fp32_mfma_157TF"] = round(useful / t / 157.3e12, 4)
            roof_nufft[tag] = e
    kernels_json = {}
    for k, (ms, n) in kern_avg.items():
        e = {"avg_ms": round(ms, 4), "launches": n}
        if k in ab:
            e["achieved_GBs"] = round(ab[k] / (ms * 1e-3) / 1e9, 1)
            e["frac_of_8TBs"] = round(ab[k] / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
        kernels_json[k] = e

    # time-segmented SENSE and the multi-kernel normal operator at the same
    # workload, L = 4 segments of a smooth synthetic rate map (interpolators
    # fitted once, untimed); one segment costs one batched NUFFT of all coils
    segments = None
    if "segments" in phases:
        from hastycucompute_amd.mri import NORMAL_IDT_T1_OP, SenseOffresonance
        from hastycucompute_amd.offres import make_offresonance_interpolators
        L = 4
        zz, yy, xx = torch.meshgrid(*[torch.linspace(-1, 1, N, device=dev)] * 3, indexing="ij")
        fmap = torch.complex(torch.full_like(xx, -10.0), 2 * math.pi * 60.0 * (xx ** 2 + 0.5 * yy - 0.3 * zz))
        times = torch.linspace(0.0, 5e-3, samples, dtype=torch.float64, device=dev).repeat(spokes)
        bil, cjl = make_offresonance_interpolators(fmap, times, L, 64)
        sop = SenseOffresonance(coords, smaps, bil, cjl, tol=1e-5)
        img1 = torch.empty_like(x)
        fwd = run_phase(f"time-segmented forward (L={L})", lambda: sop.forward(x, out=y))
        adj = run_phase(f"time-segmented adjoint, coil-combined (L={L})", lambda: sop.adjoint(y, combine=True, out=img1))
        del sop
        kernels = kernel.unsqueeze(0).expand((L,) + tuple(kernel.shape)).contiguous()
        idt = NORMAL_IDT_T1_OP(kernels, cjl, smaps)
        idt_t = run_phase(f"NORMAL_IDT_T1_OP (L={L})", lambda: idt.apply(x, out))
        del idt, kernels
        segments = {"segments": L, "forward_ms": round(fwd * 1e3, 3), "adjoint_combined_ms": round(adj * 1e3, 3),
                    "forward_per_segment_vs_type2": round(fwd / L / res["type2"], 3) if "type2" in res else None,
                    "adjoint_per_segment_vs_type1": round(adj / L / res["type1"], 3) if "type1" in res else None,
                    "idt_t1_ms": round(idt_t * 1e3, 3),
                    "idt_t1_per_kernel_vs_aha": round(idt_t / L / res["aha"], 3) if "aha" in res else None}

    frames = None
    if "frames" in phases:
        frames = bench_frames(dev)

    cpu = None
    cpu_nufft = None
    if not args.no_cpu_baseline and "aha" in phases:
        log("[bench] CPU baseline ...")
        cpu = cpu_baseline(N, C, smaps_np, x_np, kernel.cpu().numpy(), max(1, min(args.cpu_sample_coils, C)),
                           max(1, args.cpu_reps))
        log(f"[bench] CPU baseline {cpu['value']:.4f} applies/s")
    if not args.no_cpu_baseline and ("type1" in phases or "type2" in phases):
        log("[bench] CPU NUFFT baseline ...")
        cpu_nufft = cpu_baseline_nufft(N, C, coords_np, smaps_np, x_np, max(1, min(args.cpu_nufft_coils, C)))
        log(f"[bench] CPU NUFFT baseline {cpu_nufft}")

    line = {
        "metric": "A^H A applies/s + NUFFT Mpts/s, 128^3x32 coils, <=2e-5 rel err vs FP64",
        "value": round(value, 3) if value else None,
        "unit": "applies/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(aha_t * 1e3, 4) if aha_t else None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 (fp32 arithmetic; fp64 CG scalars and PSF precompute)",
        "data": "synthetic (complex Gaussian image, Gaussian-bump coil maps, golden-means kooshball)",
        "config": {"workload": args.workload, "image": [N, N, N], "coils": C, "points_per_coil": M,
                   "spokes": spokes, "samples": samples, "toeplitz_grid": [2 * N] * 3, "eps": 1e-5,
                   "kb_width": p2.info()["width"], "fine_grid": list(p2.info()["nf"])[-3:]},
        "nufft_type2_mpts_s": round(C * M / res["type2"] / 1e6, 1) if "type2" in res else None,
        "nufft_type1_mpts_s": round(C * M / res["type1"] / 1e6, 1) if "type1" in res else None,
        "plain_per_coil_gpu": ({
            "aha_applies_s": round(1.0 / plain["aha"], 3),
            "nufft_type2_mpts_s": round(C * M / plain["type2"] / 1e6, 1),
            "nufft_type1_mpts_s": round(C * M / plain["type1"] / 1e6, 1),
            "batched_speedup": {k: round(plain[k] / res[k], 3) for k in plain if k in res},
        } if plain else None),
        "batched_frames_hipgraph": frames,
        "time_segmented_and_multi_kernel": segments,
        "aha_pipeline_roofline": ({"achieved": round(ab["aha"] / aha_t / 1e9, 1), "peak": HBM_PEAK_GBS,
                                   "unit": "GB/s", "frac": round(ab["aha"] / aha_t / 1e9 / HBM_PEAK_GBS, 4),
                                   "algorithmic_bytes": ab["aha"]} if aha_t else None),
        "roofline": roof,
        "roofline_nufft": roof_nufft or None,
        "kernels": kernels_json,
        "accuracy_rel_l2": acc,
        "cg30": cg30,
        "psf_precompute_s": round(psf_s, 3),
        "setpts_s": round(setpts_s, 3),
        "cpu_baseline": cpu,
        "cpu_baseline_nufft": cpu_nufft,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
