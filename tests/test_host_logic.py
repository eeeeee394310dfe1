"""CPU: host-side logic of the Python mirror (validation, enums, synthetic
inputs) -- nothing here launches a kernel."""
import math

import numpy as np
import pytest
import torch

from hastycucompute_amd import fft as hf
from hastycucompute_amd import synth


def test_enums_mirror_reference():
    # nufft.cppm:42-70, toeplitz.cppm:47-56
    assert hf.nufft_type.FORWARD == hf.nufft_type.TYPE_2 == 2
    assert hf.nufft_type.BACKWARD == hf.nufft_type.TYPE_1 == 1
    assert hf.nufft_sign.DEFAULT_TYPE_1 == 1 and hf.nufft_sign.DEFAULT_TYPE_2 == -1
    assert [int(t) for t in hf.ToeplitzMultType] == [0, 1, 2]
    assert [int(t) for t in hf.ToeplitzAccumulateType] == [0, 1]


def test_verify_coords_messages():
    a = torch.zeros(4)
    b = torch.zeros(5)
    with pytest.raises(RuntimeError, match="x and y"):
        hf.verify_coords([a, b])
    with pytest.raises(RuntimeError, match="x,y"):
        hf.verify_coords([a, a, b])
    hf.verify_coords([a, a, a])


def test_plan_rejects_type3_and_bad_dim_before_gpu():
    with pytest.raises(RuntimeError):
        hf.nufft_plan.make(hf.nufft_opts(nmodes=(8, 8, 8)), hf.nufft_type.TYPE_3)
    with pytest.raises(RuntimeError):
        hf.nufft_plan.make(hf.nufft_opts(nmodes=(8,)), hf.nufft_type.TYPE_1)
    with pytest.raises(RuntimeError):
        hf.nufft_plan.make(hf.nufft_opts(nmodes=(8, 8)), hf.nufft_type.TYPE_1, torch.float16)


def test_cpu_tensors_are_refused():
    with pytest.raises(RuntimeError, match="GPU"):
        hf.toeplitz_multiply(torch.zeros((1, 4, 4, 4), dtype=torch.complex64),
                             torch.zeros((8, 8, 8), dtype=torch.complex64))


def test_kb_width_rule():
    assert hf.kb_width(1e-5) == 6
    assert hf.kb_width(1e-6) == 7
    assert hf.kb_width(1e-9) == 10


def test_kb_width_library_matches_oracle():
    """The library's width rule (host arithmetic, no GPU) is the oracle's, at
    both upsampling factors, 2-D and 3-D; unreachable tolerances give 17."""
    from oracle.kb import kernel_width
    for sigma in (2.0, 1.25, 1.5):
        for dim in (2, 3):
            for tol in (1e-2, 1e-3, 3e-4, 1e-5, 1e-6, 1e-8, 1e-10):
                assert hf.kb_width(tol, sigma, dim) == kernel_width(tol, sigma, dim), (tol, sigma, dim)
                assert hf.kb_width(tol, sigma, dim, torch.float64) == kernel_width(tol, sigma, dim, "f64")
    assert hf.kb_width(1e-14, 1.25, 3, torch.float64) == 17
    assert hf.kb_width(1e-5, 1.25, 3) == 17 and hf.kb_width(1e-5, 1.25, 3, torch.float64) == 9


def test_kooshball_geometry():
    kz, ky, kx = synth.kooshball(64, 32)
    for k in (kz, ky, kx):
        assert k.dtype == np.float32 and k.min() >= -np.pi - 1e-6 and k.max() <= np.pi + 1e-6
    r = np.sqrt(kz.astype(np.float64) ** 2 + ky ** 2 + kx ** 2).reshape(64, 32)
    expect = np.abs(-np.pi + 2 * np.pi * np.arange(32) / 32)
    assert np.allclose(r, expect[None], atol=1e-5)
    # golden-means directions: cos(theta) = frac(s * 0.4656)
    s = 5
    d = np.array([kz[s * 32], ky[s * 32], kx[s * 32]]) / -np.pi
    assert abs(d[0] - math.fmod(s * synth.PHI1, 1.0)) < 1e-6


def test_coil_maps_rss_normalised():
    S = synth.coil_maps((12, 10, 8), 4)
    rss = np.sqrt((np.abs(S) ** 2).sum(0))
    assert np.allclose(rss, 1.0, atol=1e-5)


def test_spiral_in_range():
    ky, kx = synth.spiral2d(4, 256)
    assert np.all(np.abs(ky) <= np.pi) and np.all(np.abs(kx) <= np.pi)


def test_stack_of_spirals_geometry():
    kz, ky, kx = synth.stack_of_spirals(40, 24, 2048)
    assert kz.size == ky.size == kx.size == 40 * 24 * 2048
    planes = np.unique(kz)
    assert planes.size == 40 and np.allclose(np.diff(planes), 2 * np.pi / 40, atol=1e-5)
    assert np.all(np.abs(ky) <= np.pi) and np.all(np.abs(kx) <= np.pi)
    # every partition carries the same in-plane spiral set
    assert np.array_equal(ky[:24 * 2048], ky[24 * 2048:2 * 24 * 2048])


def test_ellipse_phantom_is_deterministic():
    a = synth.ellipse_phantom((64, 48), seed=3)
    b = synth.ellipse_phantom((64, 48), seed=3)
    assert a.dtype == np.complex64 and a.shape == (64, 48) and np.array_equal(a, b)
    assert np.abs(a).max() > 0.1


def test_trajectory_penalty_cpu():
    from hastycucompute_amd import trajectory
    # a straight line sampled uniformly has zero slew; a step in speed has
    # non-zero slew, and speeds above gmax are penalised
    t = torch.linspace(-1, 1, 64, dtype=torch.float64)
    k = [t.clone(), torch.zeros_like(t)]
    assert float(trajectory.penalty(k, 1, gmax=1.0, slew_weight=1.0, grad_weight=1.0)) < 1e-20
    bent = [torch.cat([t[:32], t[32:] * 2]), torch.zeros_like(t)]
    assert float(trajectory.penalty(bent, 1, gmax=1.0, slew_weight=1.0, grad_weight=0.0)) > 0
    fast = [t * 100, torch.zeros_like(t)]
    assert float(trajectory.penalty(fast, 1, gmax=0.5, slew_weight=0.0, grad_weight=1.0)) > 0


def test_autotune_candidates_and_options():
    from hastycucompute_amd import autotune
    for nt in (1, 2):
        for dim in (2, 3):
            c = autotune._candidates(nt, dim)
            assert c and all(len(x["tile"]) == 3 for x in c)
            if dim == 2:
                assert all(x["tile"][0] == 1 for x in c)
    opts = autotune.plan_options({"tile": [8, 8, 16], "max_subproblem": 1024, "ms": 1.0, "trials": []})
    assert opts == {"tile": (8, 8, 16), "max_subproblem": 1024}
    hf.nufft_opts(nmodes=(8, 8, 8), **opts)      # accepted by the options block


# ---------------------------------------------------------------------------
# time-segmentation interpolators (hastycucompute_amd.offres) vs the oracle
# ---------------------------------------------------------------------------
def _fmap_on_grid(rng, shape, values):
    return np.asarray(values)[rng.integers(0, len(values), size=shape)]


def test_offres_exact_for_few_values():
    """A map with L distinct values on the bin centres is fitted exactly."""
    import torch
    from hastycucompute_amd.offres import make_offresonance_interpolators
    from oracle.offres import interpolators
    rng = np.random.default_rng(0)
    nb = 11
    grid = np.linspace(-300.0, 200.0, nb)
    vals = 1j * grid[[0, 4, 10]]                       # three frequencies (rad/s), on centres
    fmap = _fmap_on_grid(rng, (6, 5, 4), vals)
    t = np.sort(rng.uniform(0.0, 8e-3, 300))
    B, Ct = make_offresonance_interpolators(torch.from_numpy(fmap), torch.from_numpy(t), 3, nb, b0_only=True,
                                            dtype=torch.complex128)
    E = np.exp(fmap.reshape(-1)[None, :] * t[:, None])
    approx = B.numpy().T @ Ct.numpy().reshape(3, -1)
    assert np.linalg.norm(approx - E) / np.linalg.norm(E) < 1e-9
    Bo, Cto = interpolators(fmap, t, 3, nb, b0_only=True)
    assert np.abs(B.numpy() - Bo).max() < 1e-8 * np.abs(Bo).max()
    assert np.abs(Ct.numpy() - Cto).max() < 1e-12


def test_offres_error_falls_with_segments_and_matches_oracle():
    import torch
    from hastycucompute_amd.offres import make_offresonance_interpolators
    from oracle.offres import interpolators
    rng = np.random.default_rng(1)
    shape = (8, 8, 8)
    zz, yy, xx = np.meshgrid(*[np.linspace(-1, 1, n) for n in shape], indexing="ij")
    fmap = -20.0 + 1j * 2 * np.pi * 60.0 * (xx ** 2 + 0.5 * yy - 0.3 * zz)   # R2* decay + smooth field
    t = np.linspace(0.0, 10e-3, 500)
    E = np.exp(fmap.reshape(-1)[None, :] * t[:, None])
    errs = []
    for L in (2, 4, 8):
        B, Ct = make_offresonance_interpolators(torch.from_numpy(fmap), torch.from_numpy(t), L, 40,
                                                dtype=torch.complex128)
        Bo, Cto = interpolators(fmap, t, L, 40)
        assert np.linalg.norm(B.numpy() - Bo) <= 1e-6 * np.linalg.norm(Bo)
        approx = B.numpy().T @ Ct.numpy().reshape(L, -1)
        errs.append(np.linalg.norm(approx - E) / np.linalg.norm(E))
    assert errs[0] > errs[1] > errs[2] and errs[2] < 1e-2, errs


def test_offres_repeated_unsorted_times_match_oracle():
    """Readout times that repeat (every spoke has the same ones) and come
    unsorted: each column of B is the fit at its own time, as the oracle
    computes it column by column."""
    import torch
    from hastycucompute_amd.offres import make_offresonance_interpolators
    from oracle.offres import interpolators
    rng = np.random.default_rng(5)
    shape = (6, 7, 5)
    fmap = -15.0 + 1j * 2 * np.pi * 40.0 * rng.uniform(-1, 1, shape)
    readout = rng.uniform(0.0, 6e-3, 37)
    t = readout[rng.integers(0, 37, 600)]               # 600 samples, 37 distinct times, unsorted
    B, Ct = make_offresonance_interpolators(torch.from_numpy(fmap), torch.from_numpy(t), 4, 30,
                                            dtype=torch.complex128)
    Bo, Cto = interpolators(fmap, t, 4, 30)
    assert B.shape == (4, 600)
    assert np.linalg.norm(B.numpy() - Bo) <= 1e-6 * np.linalg.norm(Bo)
    # equal times give equal columns
    j0, j1 = np.nonzero(t == t[0])[0][:2]
    assert np.array_equal(B.numpy()[:, j0], B.numpy()[:, j1])


def test_offres_autocorrelated_histogram():
    import torch
    from hastycucompute_amd.offres import histogram
    from oracle.offres import histogram as ohist
    rng = np.random.default_rng(2)
    fmap = 1j * rng.uniform(-50, 50, (5, 6, 7))
    zk, h = histogram(torch.from_numpy(fmap), 16, b0_only=True, autocorrelate=True)
    zko, ho = ohist(fmap, 16, b0_only=True, autocorrelate=True)
    assert h.numel() == 31 and abs(float(h.sum()) - fmap.size ** 2) < 1e-6
    assert np.allclose(h.numpy(), ho) and np.allclose(zk.numpy(), zko)
    span = fmap.imag.max() - fmap.imag.min()
    assert np.isclose(zk.imag.max().item(), span) and np.isclose(zk.imag.min().item(), -span)


def test_binary_tensor_round_trip(tmp_path):
    import torch
    from hastycucompute_amd.io import export_binary_tensor, import_binary_tensor
    g = torch.Generator().manual_seed(0)
    for dt, shape in ((torch.complex64, (3, 4, 5)), (torch.float32, (7,)), (torch.complex128, (2, 3))):
        x = torch.randn(shape, dtype=dt, generator=g)
        p = tmp_path / f"blk_{str(dt)[6:]}.htc"
        export_binary_tensor(x, p)
        assert p.stat().st_size == x.numel() * x.element_size()   # raw bytes, no header
        y = import_binary_tensor(p, shape, dt)
        assert torch.equal(x, y)
    # the bytes are the C-order elements: a reinterpretation with another shape sees them in order
    z = import_binary_tensor(tmp_path / "blk_float32.htc", (7,), torch.float32)
    assert z.shape == (7,)
    with pytest.raises(RuntimeError, match="Exception opening/reading file"):
        import_binary_tensor(p, (4, 3), torch.complex128)
    with pytest.raises(RuntimeError, match="Could not open file for reading"):
        import_binary_tensor(tmp_path / "missing.htc", (1,), torch.float32)


def test_torch_library_ops_registered_with_fake_kernels():
    import torch
    from torch._subclasses.fake_tensor import FakeTensorMode
    import hastycucompute_amd.ops  # noqa: F401
    for name in ("nufft_type2", "nufft_type1", "toeplitz_normal", "sense_cg"):
        assert hasattr(torch.ops.hasty, name)
    with FakeTensorMode():
        img = torch.empty(3, 8, 6, 4, dtype=torch.complex64)
        k = [torch.empty(100) for _ in range(3)]
        assert torch.ops.hasty.nufft_type2(img, k, 1e-5).shape == (3, 100)
        kd = torch.empty(3, 100, dtype=torch.complex64)
        assert torch.ops.hasty.nufft_type1(kd, k, [8, 6, 4], 1e-5).shape == (3, 8, 6, 4)
        x = torch.empty(8, 6, 4, dtype=torch.complex64)
        kern = torch.empty(16, 12, 8, dtype=torch.complex64)
        S = torch.empty(2, 8, 6, 4, dtype=torch.complex64)
        assert torch.ops.hasty.toeplitz_normal(x, kern, S).shape == x.shape
        assert torch.ops.hasty.sense_cg(x, kern, S, 5).shape == x.shape
