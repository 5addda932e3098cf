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

from ._lib import Weights, check, lib, rg_apply_job

__all__ = ["RegridApply", "regrid_apply", "RegridApplyViews", "regrid_apply_views"]


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


def _apply_views(weights: tuple, x: torch.Tensor, transpose: bool) -> torch.Tensor:
    """One batched launch: forward, every element's weights applied to the
    same x, outputs concatenated; transpose, element k's adjoint applied to
    its slice of x, the K results stacked (K, n)."""
    if not x.is_cuda or x.dim() != 1:
        raise ValueError("regrid_apply_views needs a 1-D device tensor")
    if x.dtype == torch.float64:
        fn = lib().rg_apply_batched_f64
    elif x.dtype == torch.float32:
        fn = lib().rg_apply_batched_f32
    else:
        raise TypeError(f"regrid_apply_views supports float64 and float32, not {x.dtype}")
    x = x.contiguous()
    K = len(weights)
    jobs = (rg_apply_job * K)()
    es = x.element_size()
    if not transpose:
        n_in = weights[0].n_in
        if any(w.n_in != n_in for w in weights) or x.numel() != n_in:
            raise ValueError("every element must take the same input cells as x has")
        y = torch.empty(sum(w.n_out for w in weights), dtype=x.dtype, device=x.device)
        off = 0
        for k, w in enumerate(weights):
            jobs[k].weights, jobs[k].transpose, jobs[k].batch = w.ptr, 0, 1
            jobs[k].values_input = x.data_ptr()
            jobs[k].values_output = y.data_ptr() + es * off
            off += w.n_out
    else:
        n_in = weights[0].n_in
        if x.numel() != sum(w.n_out for w in weights):
            raise ValueError("x must hold every element's output cells")
        y = torch.empty((K, n_in), dtype=x.dtype, device=x.device)
        off = 0
        for k, w in enumerate(weights):
            jobs[k].weights, jobs[k].transpose, jobs[k].batch = w.ptr, 1, 1
            jobs[k].values_input = x.data_ptr() + es * off
            jobs[k].values_output = y.data_ptr() + es * k * n_in
            off += w.n_out
    stream = torch.cuda.current_stream(x.device).cuda_stream
    check(fn(jobs, K, stream))
    return y


class RegridApplyViews(torch.autograd.Function):
    """y = concat_k W_k x for weights that share their input cells (one
    batched launch); backward: sum_k W_k^T g_k (one batched launch of the
    adjoints, then a sum over k)."""

    @staticmethod
    def forward(ctx, x: torch.Tensor, weights: tuple):
        ctx.weights = weights
        return _apply_views(weights, x, False)

    @staticmethod
    def backward(ctx, grad_y: torch.Tensor):
        return _apply_views(ctx.weights, grad_y.contiguous(), True).sum(0), None


def regrid_apply_views(weights, x: torch.Tensor) -> torch.Tensor:
    """Differentiable application of several device-resident weights that
    share their input grid to the same 1-D ``x``; the outputs concatenated
    in the order of ``weights``."""
    return RegridApplyViews.apply(x, tuple(weights))
