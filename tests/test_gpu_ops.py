"""GPU: the torch.library registrations (torch.ops.hasty.*) against the FP64
oracle and against the autograd Functions they share kernels with."""
import numpy as np
import pytest
import torch

from oracle.ndft import ndft_type1, ndft_type2, rel_l2
from oracle.solvers import sense_normal_direct
from tests.gpu_util import crandn, cuda, host, rand_coords, report

pytestmark = pytest.mark.gpu


def _problem(seed=0, shape=(12, 10, 8), T=3, M=2500):
    rng = np.random.default_rng(seed)
    return rng, shape, rand_coords(rng, len(shape), M), crandn(rng, (T,) + shape), crandn(rng, (T, M))


def test_nufft_ops_vs_ndft():
    import hastycucompute_amd.ops  # noqa: F401
    _, shape, coords, f, c = _problem()
    k = [cuda(a) for a in coords]
    y = torch.ops.hasty.nufft_type2(cuda(f), k, 1e-5)
    report("hasty::nufft_type2 vs NDFT", rel_l2(host(y), ndft_type2(f, coords)), 2e-5)
    img = torch.ops.hasty.nufft_type1(cuda(c), k, list(shape), 1e-5)
    report("hasty::nufft_type1 vs NDFT", rel_l2(host(img), ndft_type1(c, coords, shape)), 2e-5)


def test_nufft_op_gradients_match_autograd_functions():
    import hastycucompute_amd.ops  # noqa: F401
    from hastycucompute_amd.autograd import NufftOperator, nufft2
    rng, shape, coords, f, _ = _problem(seed=1)
    target = cuda(crandn(rng, (3, len(coords[0]))))
    grads = []
    for use_op in (True, False):
        x = cuda(f).requires_grad_(True)
        k = [cuda(a).requires_grad_(True) for a in coords]
        if use_op:
            y = torch.ops.hasty.nufft_type2(x, k, 1e-5)
        else:
            y = nufft2(x, NufftOperator(shape, ntransf=3, tol=1e-5), *k)
        loss = (y - target).abs().pow(2).sum()
        loss.backward()
        grads.append([host(x.grad)] + [host(a.grad) for a in k])
    for a, b in zip(*grads):
        assert rel_l2(a, b) < 1e-6


def test_type1_op_gradient_is_type2():
    import hastycucompute_amd.ops  # noqa: F401
    rng, shape, coords, _, c = _problem(seed=2)
    k = [cuda(a) for a in coords]
    kd = cuda(c).requires_grad_(True)
    g = cuda(crandn(rng, (3,) + shape))
    img = torch.ops.hasty.nufft_type1(kd, k, list(shape), 1e-5)
    (img.conj() * g).real.sum().backward()
    # d/dc Re<img, g> = type2(g) (torch's conjugate-gradient convention)
    ref = ndft_type2(host(g), coords)
    report("grad of hasty::nufft_type1", rel_l2(host(kd.grad), ref), 2e-5)


def test_normal_and_cg_ops():
    import hastycucompute_amd.ops  # noqa: F401
    from hastycucompute_amd import fft as hf
    rng = np.random.default_rng(3)
    shape, C, M = (12, 12, 10), 3, 3000
    coords = rand_coords(rng, 3, M)
    S = crandn(rng, (C,) + shape)
    x = crandn(rng, shape)
    kern = torch.empty(tuple(2 * n for n in shape), dtype=torch.complex64, device="cuda")
    hf.toeplitz_kernel([cuda(a) for a in coords], kern)
    out = torch.ops.hasty.toeplitz_normal(cuda(x), kern, cuda(S))
    report("hasty::toeplitz_normal vs NDFT", rel_l2(host(out), sense_normal_direct(x, S, coords)), 2e-5)
    b = out
    xs = torch.ops.hasty.sense_cg(b, kern, cuda(S), 40)
    res = torch.ops.hasty.toeplitz_normal(xs, kern, cuda(S)) - b
    r = float(torch.linalg.vector_norm(res) / torch.linalg.vector_norm(b))
    print(f"[ops] sense_cg 40 iterations: relative residual {r:.2e}", flush=True)
    assert r < 1e-2


def test_opcheck_registrations():
    import hastycucompute_amd.ops  # noqa: F401
    _, shape, coords, f, c = _problem(seed=4, shape=(8, 8, 8), M=500)
    k = [cuda(a) for a in coords]
    tests = ("test_schema", "test_autograd_registration", "test_faketensor")
    torch.library.opcheck(torch.ops.hasty.nufft_type2.default, (cuda(f), k, 1e-5), test_utils=tests)
    torch.library.opcheck(torch.ops.hasty.nufft_type1.default, (cuda(c), k, [8, 8, 8], 1e-5), test_utils=tests)
