"""GPU parity: NUFFT type 1 / type 2 through the C-ABI vs the FP64 oracle.

Bar (from the path's accuracy contract): relative L2 <= 2e-5 against the FP64
direct NDFT at requested eps 1e-5; the error is printed per result.
"""
import numpy as np
import pytest
import torch

from oracle.kb import KBPlan
from oracle.ndft import ndft_type1, ndft_type2, rel_l2
from tests.gpu_util import crandn, cuda, host, rand_coords, report

pytestmark = pytest.mark.gpu

EPS = 1e-5
TOL = 2e-5


def _plans(shape, ntransf=1, tol=EPS, dtype=torch.float32, **kw):
    from hastycucompute_amd import fft as hf
    p1 = hf.nufft_plan.make(hf.nufft_opts(nmodes=shape, ntransf=ntransf, tol=tol, **kw),
                            hf.nufft_type.TYPE_1, dtype)
    p2 = hf.nufft_plan.make(hf.nufft_opts(nmodes=shape, ntransf=ntransf, tol=tol, **kw),
                            hf.nufft_type.TYPE_2, dtype)
    return p1, p2


def _run(shape, coords, f, c, ntransf, **kw):
    p1, p2 = _plans(shape, ntransf, **kw)
    gco = [cuda(x) for x in coords]
    p1.setpts(gco)
    p2.setpts(gco)
    M = len(coords[0])
    out2 = torch.empty((ntransf, M), dtype=torch.complex64, device="cuda")
    p2.execute(cuda(f.reshape((ntransf,) + shape)), out2)
    out1 = torch.empty((ntransf,) + shape, dtype=torch.complex64, device="cuda")
    p1.execute(cuda(c.reshape(ntransf, M)), out1)
    return host(out1), host(out2)


@pytest.mark.parametrize("shape,M", [((16, 20, 12), 3000), ((24, 24, 24), 5000), ((7, 9, 5), 1500),
                                     ((64, 48), 4000), ((33, 20), 2500)])
def test_type1_type2_vs_ndft(shape, M):
    rng = np.random.default_rng(hash(shape) % 2**32)
    coords = rand_coords(rng, len(shape), M)
    f = crandn(rng, shape)
    c = crandn(rng, M)
    g1, g2 = _run(shape, coords, f, c, 1)
    report(f"type2 {shape} M={M}", rel_l2(g2[0], ndft_type2(f, coords)), TOL)
    report(f"type1 {shape} M={M}", rel_l2(g1[0], ndft_type1(c, coords, shape)), TOL)


def test_batched_transforms_match_single():
    rng = np.random.default_rng(5)
    shape, M, T = (20, 16, 24), 4000, 3
    coords = rand_coords(rng, 3, M)
    f = crandn(rng, (T,) + shape)
    c = crandn(rng, (T, M))
    g1, g2 = _run(shape, coords, f, c, T)
    for t in range(T):
        report(f"batched type2 t={t}", rel_l2(g2[t], ndft_type2(f[t], coords)), TOL)
        report(f"batched type1 t={t}", rel_l2(g1[t], ndft_type1(c[t], coords, shape)), TOL)


def test_group_and_batch_options():
    """Every (coil_group, max_batch, tile) option gives the same transform."""
    rng = np.random.default_rng(6)
    shape, M, T = (16, 16, 16), 3000, 5
    coords = rand_coords(rng, 3, M)
    f = crandn(rng, (T,) + shape)
    c = crandn(rng, (T, M))
    ref1 = ndft_type1(c, coords, shape)
    ref2 = ndft_type2(f, coords)
    for kw in ({"coil_group": 1}, {"coil_group": 3, "max_batch": 2}, {"tile": (8, 8, 8)},
               {"max_subproblem": 64}):
        g1, g2 = _run(shape, coords, f, c, T, **kw)
        report(f"type2 opts={kw}", rel_l2(g2, ref2), TOL)
        report(f"type1 opts={kw}", rel_l2(g1, ref1), TOL)


def test_kooshball_vs_kb_oracle_and_ndft_subset():
    """64^3 radial kooshball (1024 spokes x 128): full output vs the FP64 KB
    restatement, and a point subset vs the direct NDFT."""
    from hastycucompute_amd import synth
    shape = (64, 64, 64)
    coords = synth.kooshball(1024, 128)
    M = len(coords[0])
    f = synth.image(shape, 0)
    rng = np.random.default_rng(9)
    c = crandn(rng, M)
    g1, g2 = _run(shape, coords, f, c, 1)
    kb = KBPlan(shape, EPS).setpts(coords)
    report("kooshball type2 vs KB oracle", rel_l2(g2[0], kb.type2(f)), 2e-6)
    report("kooshball type1 vs KB oracle", rel_l2(g1[0], kb.type1(c)), 2e-6)
    sel = rng.choice(M, 1500, replace=False)
    sub = [x[sel] for x in coords]
    report("kooshball type2 vs NDFT (subset)", rel_l2(g2[0][sel], ndft_type2(f, sub)), TOL)


def test_special_points():
    """k on the boundaries of [-pi, pi), at 0, and beyond (periodic)."""
    shape = (12, 10, 8)
    vals = np.array([-np.pi, np.pi - 1e-6, 0.0, 1e-7, -1e-7, 2.5, -3.0, 3 * np.pi / 2],
                    np.float32)
    rng = np.random.default_rng(3)
    coords = [rng.permutation(vals) for _ in range(3)]
    f = crandn(rng, shape)
    c = crandn(rng, len(vals))
    g1, g2 = _run(shape, coords, f, c, 1)
    report("special points type2", rel_l2(g2[0], ndft_type2(f, coords)), TOL)
    report("special points type1", rel_l2(g1[0], ndft_type1(c, coords, shape)), TOL)


def test_single_point_at_origin():
    """One point at k = 0: type 2 is sum(f), type 1 is a constant image."""
    shape = (10, 12, 8)
    rng = np.random.default_rng(4)
    f = crandn(rng, shape)
    coords = [np.zeros(1, np.float32)] * 3
    c = np.array([2.0 - 1.0j], np.complex64)
    g1, g2 = _run(shape, coords, f, c, 1)
    report("origin type2 = sum f", rel_l2(g2[0], np.array([f.astype(np.complex128).sum()])), TOL)
    report("origin type1 = const", rel_l2(g1[0], np.full(shape, c[0], np.complex128)), TOL)


def test_empty_points():
    shape = (8, 8, 8)
    g1, g2 = _run(shape, [np.zeros(0, np.float32)] * 3, np.ones(shape, np.complex64),
                  np.zeros(0, np.complex64), 1)
    assert g2.shape == (1, 0)
    assert np.all(g1 == 0)


def test_adjointness():
    rng = np.random.default_rng(11)
    shape, M = (32, 24, 20), 20000
    coords = rand_coords(rng, 3, M)
    f = crandn(rng, shape)
    c = crandn(rng, M)
    g1, g2 = _run(shape, coords, f, c, 1)
    lhs = np.vdot(g2[0].astype(np.complex128), c.astype(np.complex128))
    rhs = np.vdot(f.astype(np.complex128), g1[0].astype(np.complex128))
    err = abs(lhs - rhs) / abs(lhs)
    report("adjointness <Af,c> = <f,A^H c>", err, 2e-5)


def test_fp64_plan():
    rng = np.random.default_rng(12)
    shape, M = (12, 10, 16), 2000
    coords = [x.astype(np.float64) for x in rand_coords(rng, 3, M)]
    from hastycucompute_amd import fft as hf
    p1 = hf.nufft_plan.make(hf.nufft_opts(nmodes=shape, ntransf=1, tol=1e-10), hf.nufft_type.TYPE_1,
                            torch.float64)
    p2 = hf.nufft_plan.make(hf.nufft_opts(nmodes=shape, ntransf=1, tol=1e-10), hf.nufft_type.TYPE_2,
                            torch.float64)
    gco = [cuda(x) for x in coords]
    p1.setpts(gco)
    p2.setpts(gco)
    c = crandn(rng, (1, M), np.complex128)
    f = crandn(rng, (1,) + shape, np.complex128)
    o1 = torch.empty((1,) + shape, dtype=torch.complex128, device="cuda")
    o2 = torch.empty((1, M), dtype=torch.complex128, device="cuda")
    p1.execute(cuda(c), o1)
    p2.execute(cuda(f), o2)
    report("fp64 type1 tol 1e-10", rel_l2(host(o1)[0], ndft_type1(c[0], coords, shape)), 1e-9)
    report("fp64 type2 tol 1e-10", rel_l2(host(o2)[0], ndft_type2(f[0], coords)), 1e-9)


@pytest.mark.parametrize("spokes,kw", [(300, {}), (300, {"max_subproblem": 200}),
                                       (2400, {"max_subproblem": 3000})])
def test_paired_interpolation_is_bitwise_single(spokes, kw):
    """Type 2 on the 16^3 tile interpolates two transforms per block from
    interleaved LDS cells, after setpts reorders each subproblem's points by
    LDS bank class.  Each point's value is the same sum in the same order as
    in the one-transform kernel (an explicit coil_group keeps that kernel and
    the cell order), so the outputs agree bit for bit; odd batch (the last
    pair stages a zero partner), split subproblems with a small cap, and
    subproblems of several rounds of 1024 points with a partial last one
    (2400 spokes: thousands of points per central tile, cap 3000)."""
    from hastycucompute_amd import fft as hf
    from hastycucompute_amd import synth
    shape, T = (24, 24, 24), 5
    coords = synth.kooshball(spokes, 48)
    M = len(coords[0])
    rng = np.random.default_rng(31)
    f = crandn(rng, (T,) + shape)
    outs = []
    for opts in (dict(kw), dict(kw, coil_group=1)):
        p2 = hf.nufft_plan.make(hf.nufft_opts(nmodes=shape, ntransf=T, tol=EPS, **opts), hf.nufft_type.TYPE_2)
        p2.setpts([cuda(c) for c in coords])
        out = torch.empty((T, M), dtype=torch.complex64, device="cuda")
        p2.execute(cuda(f), out)
        outs.append(host(out))
    assert np.array_equal(outs[0], outs[1])
    sel = rng.choice(M, 1000, replace=False)
    sub = [x[sel] for x in coords]
    for t in (0, T - 1):
        report(f"paired type2 {kw} t={t} vs NDFT (subset)", rel_l2(outs[0][t][sel], ndft_type2(f[t], sub)), TOL)


def test_fp64_plan_standard_tile():
    """A complex128 type-2 plan at eps 1e-5 on the 16^3 tile (its 16-B cells
    take the bank-class point order) against the FP64 NDFT."""
    from hastycucompute_amd import fft as hf
    rng = np.random.default_rng(13)
    shape, M, T = (24, 24, 24), 3000, 2
    coords = [x.astype(np.float64) for x in rand_coords(rng, 3, M)]
    p2 = hf.nufft_plan.make(hf.nufft_opts(nmodes=shape, ntransf=T, tol=EPS), hf.nufft_type.TYPE_2, torch.float64)
    p2.setpts([cuda(x) for x in coords])
    f = crandn(rng, (T,) + shape, np.complex128)
    out = torch.empty((T, M), dtype=torch.complex128, device="cuda")
    p2.execute(cuda(f), out)
    g = host(out)
    for t in range(T):
        report(f"fp64 type2 16^3 tile t={t}", rel_l2(g[t], ndft_type2(f[t], coords)), TOL)


def test_plan_misuse_raises():
    from hastycucompute_amd import fft as hf
    p = hf.nufft_plan.make(hf.nufft_opts(nmodes=(8, 8, 8)), hf.nufft_type.TYPE_2)
    with pytest.raises(RuntimeError):
        p.execute(torch.zeros((1, 8, 8, 8), dtype=torch.complex64, device="cuda"),
                  torch.zeros((1, 4), dtype=torch.complex64, device="cuda"))
    with pytest.raises(RuntimeError):
        p.setpts([torch.zeros(4, device="cuda"), torch.zeros(5, device="cuda"),
                  torch.zeros(4, device="cuda")])
    p.setpts([torch.zeros(4, device="cuda")] * 3)
    with pytest.raises(RuntimeError):   # wrong output shape
        p.execute(torch.zeros((1, 8, 8, 8), dtype=torch.complex64, device="cuda"),
                  torch.zeros((1, 5), dtype=torch.complex64, device="cuda"))
    p.free()
    with pytest.raises(RuntimeError):
        p.setpts([torch.zeros(4, device="cuda")] * 3)


@pytest.mark.parametrize("ntransf,kw", [(32, {}), (8, {"max_subproblem": 128}), (3, {"coil_group": 2}),
                                        (32, {"tile": (4, 4, 4)})])
def test_many_coil_kooshball_vs_kb_oracle(ntransf, kw):
    """Radial sampling piles points onto the k-space centre: exercises split
    tiles and the (transform, re/im) lane mapping of the spreading kernel."""
    from hastycucompute_amd import synth
    shape = (32, 32, 32)
    coords = synth.kooshball(256, 64)
    M = len(coords[0])
    rng = np.random.default_rng(21)
    f = crandn(rng, (ntransf,) + shape)
    c = crandn(rng, (ntransf, M))
    g1, g2 = _run(shape, coords, f, c, ntransf, **kw)
    kb = KBPlan(shape, EPS).setpts(coords)
    report(f"kooshball C={ntransf} {kw} type2 vs KB oracle", rel_l2(g2, kb.type2(f)), 2e-6)
    report(f"kooshball C={ntransf} {kw} type1 vs KB oracle", rel_l2(g1, kb.type1(c)), 2e-6)


@pytest.mark.parametrize("ntransf,kw", [(10, {}), (20, {}), (34, {}), (40, {"max_batch": 18}), (17, {})])
def test_spread_partial_transform_groups(ntransf, kw):
    """3-D complex64 spreading takes the plan's transform group (the next power
    of two covering the batch, at most 32) per block, on the MFMA kernel for
    groups of 16 or 32 and even batches.  Batches that leave a group partly
    filled (10 of 16, 20 of 32, 34 = 32 + 2, chunks of 18, 18 and 4 of 32)
    read whole value rows and store only their own transforms: each transform
    equals, bit for bit, the same transform in a full plan of the same group
    (16 or 32 transforms, independent of one another), and all match the KB
    oracle.
    An odd batch (17) runs the cell-per-lane kernel: oracle only."""
    from hastycucompute_amd import fft as hf, synth
    shape = (32, 32, 32)
    coords = synth.kooshball(256, 64)
    M = len(coords[0])
    gco = [cuda(x) for x in coords]
    rng = np.random.default_rng(ntransf)
    c = crandn(rng, (ntransf, M))
    print(f"[case] partial transform groups ntransf={ntransf} {kw}", flush=True)

    def type1(vals, T, **opts):
        p = hf.nufft_plan.make(hf.nufft_opts(nmodes=shape, ntransf=T, tol=EPS, **opts), hf.nufft_type.TYPE_1)
        p.setpts(gco)
        out = torch.empty((T,) + shape, dtype=torch.complex64, device="cuda")
        p.execute(cuda(vals), out)
        p.free()
        return out

    got = type1(c, ntransf, **kw)
    if ntransf % 2 == 0:
        G = 16 if ntransf <= 16 else 32
        for lo in range(0, ntransf, G):
            hi = min(ntransf, lo + G)
            full = np.zeros((G, M), np.complex64)
            full[:hi - lo] = c[lo:hi]
            ref = type1(full, G)
            assert torch.equal(got[lo:hi], ref[:hi - lo]), f"transforms {lo}..{hi - 1} differ from the {G}-transform plan"
    kb = KBPlan(shape, EPS).setpts(coords)
    report(f"partial groups C={ntransf} {kw} type1 vs KB oracle", rel_l2(host(got), kb.type1(c)), 2e-6)


def test_setpts_is_deterministic():
    """The sorted point order and the visit lists come from stable sorts, so
    the same trajectory gives bitwise the same transforms on every setpts."""
    from hastycucompute_amd import synth
    shape, C = (20, 24, 16), 4
    k = [cuda(c) for c in synth.kooshball(400, 40)]
    M = int(k[0].shape[0])
    rng = np.random.default_rng(31)
    c = cuda(crandn(rng, (C, M)))
    f = cuda(crandn(rng, (C,) + shape))
    outs1, outs2 = [], []
    for _ in range(3):
        p1, p2 = _plans(shape, C)
        p1.setpts(k)
        p2.setpts(k)
        o1 = torch.empty((C,) + shape, dtype=torch.complex64, device="cuda")
        o2 = torch.empty((C, M), dtype=torch.complex64, device="cuda")
        p1.execute(c, o1)
        p2.execute(f, o2)
        p1.setpts(k)            # and again on the same plan
        o1b = torch.empty_like(o1)
        p1.execute(c, o1b)
        assert torch.equal(o1, o1b)
        outs1.append(o1)
        outs2.append(o2)
    assert all(torch.equal(outs1[0], o) for o in outs1[1:])
    assert all(torch.equal(outs2[0], o) for o in outs2[1:])
    print("[parity] setpts: bitwise identical type 1 / type 2 over repeated setpts and plans", flush=True)


@pytest.mark.parametrize("sign,ntransf", [(-1, 2), (1, 3), (1, 1)])
def test_fused_fine_grid_xy_pass_128_planes(sign, ntransf):
    """With 128 x 128 planes on a 256 x 256 fine grid, type 2 runs its x and y
    passes as one kernel (y split by residue classes, x in natural order, the
    centred-mode shift folded in) and type 1 its y and x inverse passes: vs
    FP64 NDFT, both signs, one and several transforms."""
    shape = (4, 128, 128)
    rng = np.random.default_rng(21 + ntransf)
    M = 1500
    coords = rand_coords(rng, 3, M)
    f = crandn(rng, (ntransf,) + shape)
    c = crandn(rng, (ntransf, M))
    out1, out2 = _run(shape, coords, f, c, ntransf, sign=sign)
    report(f"fused fine-grid type 2 {shape} sign={sign} T={ntransf}",
           rel_l2(out2, ndft_type2(f, coords, sign=sign)), TOL)
    report(f"type 1 {shape} sign={sign}", rel_l2(out1, ndft_type1(c, coords, shape, sign=sign)), TOL)


def test_fused_fine_grid_sense_forward_128_planes():
    """SENSE forward on 128 x 128 planes: maps and image ride in the fused
    pass's load diagonals (as in the row pass) vs the FP64 NDFT model."""
    from hastycucompute_amd import mri
    from oracle.solvers import sense_forward
    shape = (3, 128, 128)
    rng = np.random.default_rng(5)
    coords = rand_coords(rng, 3, 1200)
    S = crandn(rng, (3,) + shape)
    x = crandn(rng, shape)
    bil = np.ones((1, 1200), np.complex64)
    cjl = np.ones((1,) + shape, np.complex64)
    op = mri.SenseOffresonance([cuda(k) for k in coords], cuda(S), cuda(bil), cuda(cjl))
    y = host(op.forward(cuda(x)))
    report("fused fine-grid SENSE forward (3, 128, 128) C=3", rel_l2(y, sense_forward(x, S, coords)), TOL)


@pytest.mark.parametrize("combine", [False, True])
def test_fused_fine_grid_sense_adjoint_128_planes(combine):
    """Time-segmented SENSE adjoint (L = 2) on 128 x 128 planes.  Per coil,
    the type-1 y and x passes run as one fused kernel whose epilogue applies
    conj(S_c) and conj(cjl_l) as store diagonals and accumulates over the
    segments; coil-combined, the batch sum keeps the separate passes.  Both
    vs the FP64 NDFT model."""
    from hastycucompute_amd import mri
    from oracle.normal_ops import sense_adjoint_offres
    shape, C, L, M = (3, 128, 128), 3, 2, 1200
    rng = np.random.default_rng(6)
    coords = rand_coords(rng, 3, M)
    S = crandn(rng, (C,) + shape)
    y = crandn(rng, (C, M))
    bil = crandn(rng, (L, M))
    cjl = crandn(rng, (L,) + shape)
    print(f"[case] fused fine-grid SENSE adjoint {shape} C={C} L={L} combine={combine}", flush=True)
    op = mri.SenseOffresonance([cuda(k) for k in coords], cuda(S), cuda(bil), cuda(cjl))
    got = host(op.adjoint(cuda(y), combine=combine))
    ref = sense_adjoint_offres(y, S, coords, bil, cjl)
    if combine:
        ref = ref.sum(0)
    report(f"fused fine-grid SENSE adjoint {shape} C={C} L={L} combine={combine}", rel_l2(got, ref), TOL)


@pytest.mark.parametrize("tol,dtype", [(1e-3, torch.float32), (1e-4, torch.float32), (1e-5, torch.float64)])
@pytest.mark.parametrize("shape,M,ntransf", [((24, 20, 16), 3000, 1), ((20, 16, 24), 2500, 3),
                                             ((48, 40), 3000, 2)])
def test_upsamp_1_25_meets_tolerance(tol, dtype, shape, M, ntransf):
    """nufft_upsamp_cuda::UPSAMP_1_25 (``nufft.cppm:141-150``): on a 1.25x
    fine grid the window is widened until its error estimate (aliasing plus
    the plan's rounding floor) meets tol -- w = 6 at 1e-3, 7-8 at 1e-4 in
    complex64, 9 at 1e-5 in complex128; the plan reports that width, the
    oracle's rule gives the same one, and both transforms meet 2 tol against
    the FP64 NDFT."""
    from hastycucompute_amd import fft as hf
    from oracle.kb import kernel_width, next_smooth
    d = len(shape)
    f64 = dtype == torch.float64
    rdt, cdt = (np.float64, np.complex128) if f64 else (np.float32, np.complex64)
    tdt = torch.complex128 if f64 else torch.complex64
    rng = np.random.default_rng(int(tol * 1e6) + M + ntransf)
    coords = [x.astype(rdt) for x in rand_coords(rng, d, M)]
    f = crandn(rng, (ntransf,) + shape, cdt)
    c = crandn(rng, (ntransf, M), cdt)
    p1, p2 = _plans(shape, ntransf, tol=tol, dtype=dtype, upsamp=hf.nufft_upsamp_cuda.UPSAMP_1_25)
    w = p2.info()["width"]
    assert w == p1.info()["width"] == kernel_width(tol, 1.25, d, "f64" if f64 else "f32") \
        == hf.kb_width(tol, 1.25, d, dtype)
    assert w > hf.kb_width(tol, 2.0, d, dtype), "sigma 1.25 must widen the sigma-2 window"
    nf = p2.info()["nf"][-d:]
    assert tuple(nf) == tuple(next_smooth(max(int(np.ceil(1.25 * n)), 2 * w)) for n in shape)
    gco = [cuda(x) for x in coords]
    p1.setpts(gco)
    p2.setpts(gco)
    o2 = torch.empty((ntransf, M), dtype=tdt, device="cuda")
    o1 = torch.empty((ntransf,) + shape, dtype=tdt, device="cuda")
    p2.execute(cuda(f), o2)
    p1.execute(cuda(c), o1)
    bar = 2 * tol
    tag = f"upsamp 1.25 {'c128' if f64 else 'c64'} tol={tol:g} w={w}"
    report(f"{tag} type2 {shape} T={ntransf}", rel_l2(host(o2), ndft_type2(f, coords)), bar)
    report(f"{tag} type1 {shape} T={ntransf}", rel_l2(host(o1), ndft_type1(c, coords, shape)), bar)


def test_upsamp_1_25_unreachable_tolerance_raises():
    """complex64 arithmetic on a 1.25x grid bottoms out near 6e-5 (the FP32
    rounding of the fine-grid FFT grows with the wide window), so a
    complex64 plan at 1e-5 is refused rather than missing tol silently; a
    complex128 plan at 1e-5 is made."""
    from hastycucompute_amd import fft as hf
    with pytest.raises(RuntimeError, match="kernel wider"):
        _plans((16, 16, 16), 1, tol=1e-5, upsamp=hf.nufft_upsamp_cuda.UPSAMP_1_25)
    p1, p2 = _plans((16, 16, 16), 1, tol=1e-5, dtype=torch.float64, upsamp=hf.nufft_upsamp_cuda.UPSAMP_1_25)
    assert p1.info()["width"] == p2.info()["width"] == 9


@pytest.mark.parametrize("method", ["GM_NO_SORT", "GM_SORT"])
@pytest.mark.parametrize("shape,M,ntransf", [((16, 20, 12), 3000, 1), ((20, 16, 24), 2500, 3),
                                             ((64, 48), 4000, 1), ((33, 20), 2500, 2)])
def test_global_memory_methods_vs_ndft(method, shape, M, ntransf):
    """nufft_method_cuda::GM_NO_SORT / GM_SORT (``nufft.cppm:119-130``) run the
    global-memory kernels (one thread per point; type 1 with global atomics):
    both types, 2-D and 3-D, one and several transforms, vs the FP64 NDFT."""
    from hastycucompute_amd import fft as hf
    rng = np.random.default_rng(M + ntransf + len(shape))
    coords = rand_coords(rng, len(shape), M)
    f = crandn(rng, (ntransf,) + shape)
    c = crandn(rng, (ntransf, M))
    g1, g2 = _run(shape, coords, f, c, ntransf, method=hf.nufft_method_cuda[method])
    report(f"{method} type2 {shape} T={ntransf}", rel_l2(g2, ndft_type2(f, coords)), TOL)
    report(f"{method} type1 {shape} T={ntransf}", rel_l2(g1, ndft_type1(c, coords, shape)), TOL)


def test_global_memory_method_kooshball_vs_kb_oracle():
    """GM_SORT on a kooshball (points piled on the centre, so many atomics meet
    on one cell) agrees with the LDS-tiled default to the KB oracle's bar."""
    from hastycucompute_amd import fft as hf
    from hastycucompute_amd import synth
    shape, T = (32, 32, 32), 4
    coords = synth.kooshball(256, 64)
    M = len(coords[0])
    rng = np.random.default_rng(22)
    f = crandn(rng, (T,) + shape)
    c = crandn(rng, (T, M))
    g1, g2 = _run(shape, coords, f, c, T, method=hf.nufft_method_cuda.GM_SORT)
    kb = KBPlan(shape, EPS).setpts(coords)
    report("GM_SORT kooshball type2 vs KB oracle", rel_l2(g2, kb.type2(f)), 2e-6)
    report("GM_SORT kooshball type1 vs KB oracle", rel_l2(g1, kb.type1(c)), 2e-6)
