"""GPU: gradients of the NUFFT / SENSE operators (hastycucompute_amd.autograd).

References, all in FP64:
* the exact trajectory derivative d/dk_a type2(h) = type2(-i n_a h) by the
  direct NDFT (oracle.ndft);
* torch autograd through a dense FP64 NDFT on the CPU (the loss of the
  trajectory-design workload: ||A(k)(S x) - y||^2);
* central finite differences (h = 1e-4) of that FP64 loss along random
  trajectory directions;
* a dense FP64 solve of the SENSE normal equations for the CG gradient.
Bar: relative L2 (or relative directional error) <= 1e-4 for every
gradient at plan tolerance 1e-6, including the gradient through the SENSE-CG
solve (FP32 CG on the Toeplitz operator, 150 iterations).
"""
import numpy as np
import pytest
import torch

from oracle.ndft import ndft_type2, rel_l2
from tests.gpu_util import crandn, cuda, host, rand_coords, report

pytestmark = pytest.mark.gpu


def _centred(n):
    return np.arange(n) - n // 2


@pytest.mark.parametrize("shape,M,T", [((10, 12, 8), 1500, 2), ((40, 36), 2500, 3)])
def test_coord_grad_vs_ndft_derivative(shape, M, T):
    from hastycucompute_amd import fft as hf
    rng = np.random.default_rng(11)
    coords = rand_coords(rng, len(shape), M)
    h = crandn(rng, (T,) + shape)
    w = crandn(rng, (T, M))
    p = hf.nufft_plan.make(hf.nufft_opts(nmodes=shape, ntransf=T, tol=1e-6), hf.nufft_type.TYPE_2)
    p.setpts([cuda(c) for c in coords])
    got = [host(g) for g in p.coord_grad(cuda(h), cuda(w))]
    d = len(shape)
    for a in range(d):
        n_a = _centred(shape[a]).reshape([-1 if b == a else 1 for b in range(d)])
        ref = np.zeros(M)
        for t in range(T):
            D = ndft_type2((-1j * n_a * h[t]).astype(np.complex128), [c.astype(np.float64) for c in coords])
            ref += np.real(np.conj(w[t].astype(np.complex128)) * D)
        report(f"coord_grad {shape} axis {a} vs NDFT derivative", rel_l2(got[a], ref), 1e-4)


def test_coord_grad_fp64_plan():
    from hastycucompute_amd import fft as hf
    rng = np.random.default_rng(12)
    shape, M = (12, 10, 14), 1200
    coords = [c.astype(np.float64) for c in rand_coords(rng, 3, M)]
    h = crandn(rng, (1,) + shape, np.complex128)
    w = crandn(rng, (1, M), np.complex128)
    p = hf.nufft_plan.make(hf.nufft_opts(nmodes=shape, ntransf=1, tol=1e-11), hf.nufft_type.TYPE_2,
                           torch.float64)
    p.setpts([cuda(c) for c in coords])
    got = [host(g) for g in p.coord_grad(cuda(h), cuda(w))]
    for a in range(3):
        n_a = _centred(shape[a]).reshape([-1 if b == a else 1 for b in range(3)])
        D = ndft_type2(-1j * n_a * h[0], coords)
        report(f"coord_grad fp64 axis {a}", rel_l2(got[a], np.real(np.conj(w[0]) * D)), 1e-8)


# ---------------------------------------------------------------------------
# trajectory-design loss: ||A(k)(S x) - y||^2, 2-D, FP64 NDFT reference
# ---------------------------------------------------------------------------
def _ndft_matrix(kk, shape):
    """E[j, n] = exp(-i k_j . n) for centred n (torch complex128, differentiable in k)."""
    grids = torch.meshgrid(*[torch.arange(n, dtype=torch.float64) - n // 2 for n in shape], indexing="ij")
    ph = sum(kk[a][:, None] * grids[a].reshape(1, -1) for a in range(len(shape)))
    return torch.exp(-1j * ph)


def _ref_loss(x, S, y, kk, shape):
    E = _ndft_matrix(kk, shape)
    r = (E @ (S * x[None]).reshape(S.shape[0], -1).T).T - y
    return (r.abs() ** 2).sum()


def _problem(seed, shape=(32, 32), C=4, M=2000):
    rng = np.random.default_rng(seed)
    coords = rand_coords(rng, 2, M, -3.0, 3.0)
    x = crandn(rng, shape)
    S = crandn(rng, (C,) + shape) * 0.5
    y = crandn(rng, (C, M)) * 4.0
    return coords, x, S, y


@pytest.mark.parametrize("shape,M", [((32, 32), 2000), ((64, 64), 4096)])
def test_autograd_nufft2_vs_fp64_ndft_and_fd(shape, M):
    """The trajectory-design loss at 32^2 and at 64^2 (the size the
    differentiable-NUFFT workload checks against autograd through an FP64
    NDFT): image and trajectory gradients vs FP64 autograd, trajectory
    gradient vs FP64 central differences (h = 1e-4)."""
    from hastycucompute_amd import autograd as ag
    C = 4
    coords, x, S, y = _problem(13, shape, C, M)
    op = ag.NufftOperator(shape, ntransf=C, tol=1e-6)
    k = [cuda(c).requires_grad_(True) for c in coords]
    xg = cuda(x).requires_grad_(True)
    loss = (ag.sense_forward(xg, cuda(S), op, *k) - cuda(y)).abs().pow(2).sum()
    loss.backward()
    # autograd through the FP64 NDFT
    kk = [torch.from_numpy(c.astype(np.float64)).requires_grad_(True) for c in coords]
    x64 = torch.from_numpy(x.astype(np.complex128)).requires_grad_(True)
    S64, y64 = torch.from_numpy(S.astype(np.complex128)), torch.from_numpy(y.astype(np.complex128))
    L = _ref_loss(x64, S64, y64, kk, shape)
    L.backward()
    lv, Lv = float(loss.detach()), float(L.detach())
    report(f"{shape} loss vs FP64 NDFT", abs(lv - Lv) / abs(Lv), 1e-5)
    report(f"{shape} d loss / d image vs FP64 autograd", rel_l2(host(xg.grad), x64.grad.numpy()), 1e-4)
    for a in range(2):
        report(f"{shape} d loss / d k[{a}] vs FP64 autograd", rel_l2(host(k[a].grad), kk[a].grad.numpy()), 1e-4)
    # central finite differences of the FP64 loss along random directions
    rng = np.random.default_rng(99)
    g = np.concatenate([host(k[a].grad) for a in range(2)]).astype(np.float64)
    errs = []
    with torch.no_grad():
        for _ in range(8):
            dvec = rng.standard_normal(g.size)
            dvec /= np.linalg.norm(dvec)
            hstep = 1e-4
            dk = [torch.from_numpy(dvec[a * len(coords[0]):(a + 1) * len(coords[0])]) for a in range(2)]
            kp = [kk[a].detach() + hstep * dk[a] for a in range(2)]
            km = [kk[a].detach() - hstep * dk[a] for a in range(2)]
            fd = (float(_ref_loss(x64.detach(), S64, y64, kp, shape)) -
                  float(_ref_loss(x64.detach(), S64, y64, km, shape))) / (2 * hstep)
            errs.append(abs(float(g @ dvec) - fd) / np.linalg.norm(g))
    report(f"{shape} d loss / d k vs central FD (8 directions, max rel)", max(errs), 1e-4)


def test_autograd_nufft1_vs_fp64():
    from hastycucompute_amd import autograd as ag
    shape, C, M = (24, 20), 3, 1500
    rng = np.random.default_rng(14)
    coords = rand_coords(rng, 2, M)
    c = crandn(rng, (C, M))
    target = crandn(rng, (C,) + shape) * 10
    op = ag.NufftOperator(shape, ntransf=C, tol=1e-6)
    k = [cuda(v).requires_grad_(True) for v in coords]
    cg_ = cuda(c).requires_grad_(True)
    loss = (ag.nufft1(cg_, op, *k) - cuda(target)).abs().pow(2).sum()
    loss.backward()
    kk = [torch.from_numpy(v.astype(np.float64)).requires_grad_(True) for v in coords]
    c64 = torch.from_numpy(c.astype(np.complex128)).requires_grad_(True)
    E = _ndft_matrix(kk, shape)                       # [M, n]
    img = (c64 @ E.conj()).reshape((C,) + shape)      # sum_j c_j exp(+i k_j n)
    L = (img - torch.from_numpy(target.astype(np.complex128))).abs().pow(2).sum()
    L.backward()
    report("type-1 d loss / d kdata vs FP64 autograd", rel_l2(host(cg_.grad), c64.grad.numpy()), 1e-4)
    for a in range(2):
        report(f"type-1 d loss / d k[{a}] vs FP64 autograd", rel_l2(host(k[a].grad), kk[a].grad.numpy()), 1e-4)


def test_sense_cg_implicit_gradient_vs_dense_fp64():
    from hastycucompute_amd import autograd as ag
    shape, C, M = (16, 16), 3, 700
    rng = np.random.default_rng(15)
    coords = rand_coords(rng, 2, M, -3.0, 3.0)
    S = crandn(rng, (C,) + shape) * 0.6
    x_true = crandn(rng, shape)
    y = crandn(rng, (C, M))
    op = ag.NufftOperator(shape, ntransf=C, tol=1e-6)
    k = [cuda(v).requires_grad_(True) for v in coords]
    xh = ag.sense_cg_solve(cuda(y), cuda(S), op, *k, iters=150, method="implicit")
    loss = (xh - cuda(x_true)).abs().pow(2).sum()
    loss.backward()
    # dense FP64 reference: x_hat = (sum_c A_c^H A_c)^{-1} sum_c A_c^H y_c with A_c = E diag(S_c)
    kk = [torch.from_numpy(v.astype(np.float64)).requires_grad_(True) for v in coords]
    S64 = torch.from_numpy(S.astype(np.complex128)).reshape(C, -1)
    E = _ndft_matrix(kk, shape)
    Ac = [E * S64[c][None, :] for c in range(C)]
    Nmat = sum(a.conj().T @ a for a in Ac)
    b = sum(Ac[c].conj().T @ torch.from_numpy(y[c].astype(np.complex128)) for c in range(C))
    x64 = torch.linalg.solve(Nmat, b)
    L = (x64 - torch.from_numpy(x_true.astype(np.complex128)).reshape(-1)).abs().pow(2).sum()
    L.backward()
    report("CG solution vs dense FP64 solve", rel_l2(host(xh).reshape(-1), x64.detach().numpy()), 1e-4)
    for a in range(2):
        report(f"CG-solve d loss / d k[{a}] vs dense FP64", rel_l2(host(k[a].grad), kk[a].grad.numpy()), 1e-4)


def test_truncated_cg_adjoint_gradient_vs_fp64_unrolled():
    """10 CG steps (not converged): the reverse-mode adjoint through the
    checkpointed iterates against torch FP64 autograd through the same 10
    dense-matrix CG steps."""
    from hastycucompute_amd import autograd as ag
    shape, C, M, iters = (16, 16), 3, 500, 10
    rng = np.random.default_rng(16)
    coords = rand_coords(rng, 2, M, -3.0, 3.0)
    S = crandn(rng, (C,) + shape) * 0.6
    x_true = crandn(rng, shape)
    y = crandn(rng, (C, M))
    op = ag.NufftOperator(shape, ntransf=C, tol=1e-6)
    k = [cuda(v).requires_grad_(True) for v in coords]
    yg = cuda(y).requires_grad_(True)
    xh = ag.sense_cg_solve(yg, cuda(S), op, *k, iters=iters)
    loss = (xh - cuda(x_true)).abs().pow(2).sum()
    loss.backward()
    kk = [torch.from_numpy(v.astype(np.float64)).requires_grad_(True) for v in coords]
    y64 = torch.from_numpy(y.astype(np.complex128)).requires_grad_(True)
    S64 = torch.from_numpy(S.astype(np.complex128)).reshape(C, -1)
    E = _ndft_matrix(kk, shape)
    V = float(np.prod(shape))
    Ac = [E * S64[c][None, :] for c in range(C)]
    Nmat = sum(a.conj().T @ a for a in Ac) / V
    b = sum(Ac[c].conj().T @ y64[c] for c in range(C)) / V
    x = torch.zeros_like(b)
    r = b.clone()
    p = r.clone()
    rz = torch.vdot(r, r).real
    for _ in range(iters):
        q = Nmat @ p
        alpha = rz / torch.vdot(p, q).real
        x = x + alpha * p
        r = r - alpha * q
        rz1 = torch.vdot(r, r).real
        p = r + (rz1 / rz) * p
        rz = rz1
    L = (x - torch.from_numpy(x_true.astype(np.complex128)).reshape(-1)).abs().pow(2).sum()
    L.backward()
    report(f"{iters}-step CG iterate vs FP64 unrolled CG", rel_l2(host(xh).reshape(-1), x.detach().numpy()), 1e-4)
    report("truncated-CG d loss / d y vs FP64 autograd", rel_l2(host(yg.grad), y64.grad.numpy()), 1e-4)
    for a in range(2):
        report(f"truncated-CG d loss / d k[{a}] vs FP64 autograd", rel_l2(host(k[a].grad), kk[a].grad.numpy()), 1e-4)
