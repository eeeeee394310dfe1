"""Trajectory design by gradient descent through the differentiable NUFFT.

The loss of a k-space trajectory k (interleaves x samples, 2-D or 3-D) for a
known image x with coil maps S and target data y is

    L(k) = ||A(k)(S x) - y||^2 / ||y||^2
           + slew_weight * sum |k[i+1] - 2 k[i] + k[i-1]|^2      (slew)
           + grad_weight * sum relu(|k[i+1] - k[i]| - gmax)^2    (gradient limit)

with the differences taken along each interleave.  The data term's gradient
comes from the GPU kernel-derivative interpolation; Adam updates k.
"""
from __future__ import annotations

from typing import List, Sequence

import torch

from .autograd import NufftOperator, sense_forward


def penalty(k: Sequence[torch.Tensor], ninterleaves: int, gmax: float, slew_weight: float,
            grad_weight: float) -> torch.Tensor:
    kk = torch.stack(list(k)).reshape(len(k), ninterleaves, -1)      # [d, I, S]
    d1 = kk[..., 1:] - kk[..., :-1]
    d2 = kk[..., 2:] - 2 * kk[..., 1:-1] + kk[..., :-2]
    g = torch.sqrt((d1 * d1).sum(0) + 1e-12)
    return slew_weight * (d2 * d2).sum() + grad_weight * torch.relu(g - gmax).pow(2).sum()


def design(x: torch.Tensor, smaps: torch.Tensor, k0: Sequence[torch.Tensor], y: torch.Tensor,
           ninterleaves: int, steps: int = 200, lr: float = 1e-3, gmax: float = 0.1,
           slew_weight: float = 1e-2, grad_weight: float = 1.0, tol: float = 1e-5) -> dict:
    """Runs ``steps`` Adam steps on the trajectory; returns the final
    trajectory and the loss history (data term and total)."""
    op = NufftOperator(tuple(x.shape), ntransf=int(smaps.shape[0]), tol=tol)
    k = [c.detach().clone().requires_grad_(True) for c in k0]
    opt = torch.optim.Adam(k, lr=lr)
    ynorm = float((y.abs() ** 2).sum())
    data_hist: List[float] = []
    total_hist: List[float] = []
    for _ in range(steps):
        opt.zero_grad(set_to_none=True)
        data = (sense_forward(x, smaps, op, *k) - y).abs().pow(2).sum() / ynorm
        loss = data + penalty(k, ninterleaves, gmax, slew_weight, grad_weight)
        loss.backward()
        opt.step()
        with torch.no_grad():
            for c in k:           # keep the trajectory inside [-pi, pi)
                c.clamp_(-3.14159, 3.14159)
        data_hist.append(float(data.detach()))
        total_hist.append(float(loss.detach()))
    return {"k": [c.detach() for c in k], "data_loss": data_hist, "loss": total_hist}
