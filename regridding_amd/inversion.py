"""
Iterative least-squares inversion through the autograd adjoint.

Recover an image ``x`` on one grid from observations ``y_k ~ A_k x`` on
other grids, with ``A_k`` conservative regridding weights: minimise
``sum_k 0.5 |A_k x - y_k|^2``.  The gradient comes from torch.autograd,
whose backward for each ``A_k`` is the transpose-apply kernel.  Iteration:
accelerated gradient descent (FISTA without a prox term) with the step
1 / L, L the largest eigenvalue of ``sum_k A_k^T A_k`` from power iteration.
"""

from __future__ import annotations

import torch

from .torch_ops import _apply_views, regrid_apply, regrid_apply_views

__all__ = ["lipschitz", "least_squares"]


def _fused(views):
    """The views' weights and concatenated observations when every view
    takes the same input cells (one batched launch per direction), else None."""
    ws = [w for w, _ in views]
    if len(ws) < 2 or any(w.n_in != ws[0].n_in for w in ws):
        return None
    return tuple(ws), torch.cat([y.reshape(-1) for _, y in views])


def _loss(views, x, fused=None):
    if fused is not None:
        ws, ycat = fused
        r = regrid_apply_views(ws, x) - ycat
        return 0.5 * (r * r).sum()
    total = x.new_zeros(())
    for w, y in views:
        r = regrid_apply(w, x) - y
        total = total + 0.5 * (r * r).sum()
    return total


def lipschitz(views, n: int, iters: int = 30, device=None, dtype=torch.float64) -> float:
    """Largest eigenvalue of sum_k A_k^T A_k (power iteration on the
    forward and adjoint kernels)."""
    v = torch.ones(n, dtype=dtype, device=device)
    v /= v.norm()
    lam = 0.0
    fused = _fused(views)
    for _ in range(iters):
        if fused is not None:
            ws = fused[0]
            u = _apply_views(ws, _apply_views(ws, v, False), True).sum(0)
        else:
            u = torch.zeros_like(v)
            for w, _y in views:
                u += regrid_apply(w, regrid_apply(w, v), transpose=True)
        lam = float(u.norm())
        v = u / lam
    return lam


def least_squares(views, x0: torch.Tensor, iters: int = 200, step: float | None = None):
    """``views``: sequence of (Weights, observation tensor of the output
    cells).  Returns ``(x, losses)`` -- the estimate and the loss before
    every iteration plus the final one."""
    if step is None:
        step = 1.0 / lipschitz(views, x0.numel(), device=x0.device, dtype=x0.dtype)
    x = x0.detach().clone()
    z = x.clone()
    t = 1.0
    fused = _fused(views)
    losses = []  # device scalars: no host wait inside the loop
    for _ in range(iters):
        z.requires_grad_(True)
        loss = _loss(views, z, fused)
        (g,) = torch.autograd.grad(loss, z)
        losses.append(loss.detach())
        with torch.no_grad():
            x_new = z - step * g
            t_new = 0.5 * (1.0 + (1.0 + 4.0 * t * t) ** 0.5)
            z = x_new + ((t - 1.0) / t_new) * (x_new - x)
            x, t = x_new, t_new
    with torch.no_grad():
        losses.append(_loss(views, x, fused))
    return x, [float(v) for v in torch.stack(losses).cpu()]
