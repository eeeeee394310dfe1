"""Helpers shared by the GPU parity tests."""
import numpy as np
import torch


def dev():
    return torch.device("cuda", 0)


def cuda(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.to(dev())


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def rand_coords(rng, d, M, lo=-np.pi, hi=np.pi):
    return [rng.uniform(lo, hi, M).astype(np.float32) for _ in range(d)]


def crandn(rng, shape, dtype=np.complex64):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dtype)


def report(name, err, tol):
    print(f"[parity] {name}: rel_l2 = {err:.3e} (tol {tol:.1e})", flush=True)
    assert err <= tol, f"{name}: rel_l2 {err:.3e} > {tol:.1e}"
