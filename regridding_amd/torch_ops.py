"""
PyTorch autograd wrapper of the weight application.

``RegridApply`` runs the CSR gather (``rg_apply_f64`` / ``rg_apply_f32``)
forward and, in backward, the hand-written adjoint: the same gather over the
CSR by input, i.e. the transposed weights (what the reference computes with
``regrid_from_weights(*transpose_weights(weights), ...)``,
``regridding/_weights/_weights_transposed/_weights_transposed.py:13-50``).
Both run on torch's current stream directly on the tensors' memory; PyTorch
provides device memory, streams and the autograd graph only.
"""

from __future__ import annotations

import math

import torch

from ._lib import Weights, check, lib

__all__ = ["RegridApply", "regrid_apply"]


def _apply(weights: Weights, x: torch.Tensor, transpose: bool) -> torch.Tensor:
    if not x.is_cuda:
        raise ValueError("regrid_apply needs a device tensor")
    if x.dtype == torch.float64:
        fn = lib().rg_apply_f64
    elif x.dtype == torch.float32:
        fn = lib().rg_apply_f32
    else:
        raise TypeError(f"regrid_apply supports float64 and float32, not {x.dtype}")
    n_src, n_dst = (weights.n_out, weights.n_in) if transpose else (weights.n_in, weights.n_out)
    if x.shape[-1:] != (n_src,):
        raise ValueError(f"last axis has {x.shape[-1:]} cells, the weights take {n_src}")
    x = x.contiguous()
    lead = tuple(x.shape[:-1])
    y = torch.empty(lead + (n_dst,), dtype=x.dtype, device=x.device)
    batch = math.prod(lead)
    stream = torch.cuda.current_stream(x.device).cuda_stream
    check(fn(weights.ptr, int(transpose), x.data_ptr(), y.data_ptr(), batch, stream))
    return y


class RegridApply(torch.autograd.Function):
    """y = W x over the last axis (W^T x with transpose=True); backward
    applies the other direction."""

    @staticmethod
    def forward(ctx, x: torch.Tensor, weights: Weights, transpose: bool = False):
        ctx.weights = weights
        ctx.transpose = bool(transpose)
        return _apply(weights, x, bool(transpose))

    @staticmethod
    def backward(ctx, grad_y: torch.Tensor):
        return _apply(ctx.weights, grad_y, not ctx.transpose), None, None


def regrid_apply(weights: Weights, x: torch.Tensor, transpose: bool = False) -> torch.Tensor:
    """Differentiable application of device-resident weights to ``x`` of
    shape ``(..., cells)`` (flat cell axis, input cells unless transpose)."""
    return RegridApply.apply(x, weights, transpose)
