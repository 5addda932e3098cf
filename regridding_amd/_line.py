"""
Weights along one axis: conservative 1-D and multilinear 1-D.

Host side of ``regridding/_weights/_weights_conservative.py:57-102`` (the
1-D branch) and ``regridding/_weights/_weights_multilinear.py:9-145``: the
same normalisation, axis moves and error behaviour, with the per-element
sweeps running on the GPU (``rg_conservative_1d_*``, ``rg_multilinear_1d``).

All elements of one call share ONE device handle: a block-diagonal weight
matrix whose block d relates element d's input cells to its output cells.
The per-element triples returned to the caller are registered as parts of
it, so ``regrid_from_weights`` on them applies every element in a single
launch (each output row still accumulates in the reference's order).
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _cache, _util
from ._lib import DeviceArray, Weights, check, lib

__all__ = ["conservative_1d", "multilinear_1d", "block_handle"]


def _wrap(idx: np.ndarray, n: int) -> np.ndarray:
    """numba-style wraparound of negative indices."""
    return np.where(idx < 0, idx + n, idx)


def block_handle(parts, n_in: int, n_out: int) -> Weights:
    """One device handle for per-element triples ``parts`` (each sorted by
    (input, output)): element d occupies rows d*n_in.. and columns
    d*n_out.. of a block-diagonal matrix."""
    from ._regrid import import_weights

    gi, go, gw = [], [], []
    for d, (ii, io, w) in enumerate(parts):
        gi.append(_wrap(np.asarray(ii, np.int64), n_in) + d * n_in)
        go.append(_wrap(np.asarray(io, np.int64), n_out) + d * n_out)
        gw.append(np.asarray(w, np.float64))
    D = len(parts)
    cat = (lambda xs, dt: np.concatenate(xs) if xs else np.empty(0, dt))
    return import_weights(cat(gi, np.int64), cat(go, np.int64), cat(gw, np.float64),
                          D * n_in, D * n_out)


def _register_parts(parts, handle: Weights, n_in: int, n_out: int, shape_orthogonal):
    result = np.empty(len(parts), dtype=object)
    for d, (ii, io, w) in enumerate(parts):
        _cache.register(ii, io, w, handle, part=(d, len(parts), n_in, n_out))
        result[d] = (ii, io, w)
    return result.reshape(shape_orthogonal)


def conservative_1d(x_input: np.ndarray, x_output: np.ndarray,
                    weights_input: np.ndarray | None, shape_orthogonal):
    """``x_input`` (D, n), ``x_output`` (D, m) vertices (last axis swept),
    ``weights_input`` (D, n-1) or None.  Returns the object array of
    coalesced per-element triples over ``shape_orthogonal``."""
    x_in = np.ascontiguousarray(x_input, dtype=np.float64)
    x_out = np.ascontiguousarray(x_output, dtype=np.float64)
    D, n = x_in.shape
    m = x_out.shape[1]
    L = lib()
    dxi, dxo = DeviceArray.from_host(x_in), DeviceArray.from_host(x_out)
    dwi = None
    if weights_input is not None:
        dwi = DeviceArray.from_host(np.ascontiguousarray(weights_input, dtype=np.float64))
    cells = D * (n - 1)
    counts = DeviceArray.empty((max(cells, 1),), np.int32)
    offsets = DeviceArray.empty((cells + 1,), np.int64)
    total = ctypes.c_int64()
    check(L.rg_conservative_1d_count(dxi.ptr, n, dxo.ptr, m, D, counts.ptr, offsets.ptr,
                                     ctypes.byref(total), 0))
    T = total.value
    ii = DeviceArray.empty((max(T, 1),), np.int64)
    io = DeviceArray.empty((max(T, 1),), np.int64)
    w = DeviceArray.empty((max(T, 1),), np.float64)
    check(L.rg_conservative_1d_emit(dxi.ptr, n, dxo.ptr, m, D, dwi.ptr if dwi else None,
                                    counts.ptr, offsets.ptr, ii.ptr, io.ptr, w.ptr, 0))
    off = offsets.to_host()[:: n - 1] if cells else np.zeros(D + 1, np.int64)
    hi, ho, hw = ii.to_host()[:T], io.to_host()[:T], w.to_host()[:T]
    parts = [(hi[off[d]:off[d + 1]].copy(), ho[off[d]:off[d + 1]].copy(),
              hw[off[d]:off[d + 1]].copy()) for d in range(D)]
    handle = block_handle(parts, n - 1, m - 1)
    return _register_parts(parts, handle, n - 1, m - 1, shape_orthogonal)


def multilinear_1d(indices: np.ndarray, x_input: np.ndarray, x_output: np.ndarray,
                   weights_input: np.ndarray | None, bounds: str, shape_orthogonal):
    """``indices`` (D, m) searchsorted cells of the output vertices,
    ``x_input`` (D, n), ``x_output`` (D, m), ``weights_input`` (D, n) or
    None.  Returns the object array of coalesced per-element triples."""
    x_in = np.ascontiguousarray(x_input, dtype=np.float64)
    x_out = np.ascontiguousarray(x_output, dtype=np.float64)
    idx = np.ascontiguousarray(indices, dtype=np.int64)
    D, n = x_in.shape
    m = x_out.shape[1]
    L = lib()
    dxi, dxo, didx = (DeviceArray.from_host(a) for a in (x_in, x_out, idx))
    dwi = None
    if weights_input is not None:
        dwi = DeviceArray.from_host(np.ascontiguousarray(weights_input, dtype=np.float64))
    T = 2 * D * m
    ii = DeviceArray.empty((max(T, 1),), np.int64)
    io = DeviceArray.empty((max(T, 1),), np.int64)
    w = DeviceArray.empty((max(T, 1),), np.float64)
    n_outside = ctypes.c_int64()
    check(L.rg_multilinear_1d(dxi.ptr, n, dxo.ptr, m, D, didx.ptr, dwi.ptr if dwi else None,
                              int(bounds == "nan"), ii.ptr, io.ptr, w.ptr,
                              ctypes.byref(n_outside), 0))
    if bounds == "raise" and n_outside.value:
        raise ValueError(
            f"{n_outside.value} of the output points fall outside the input "
            f"grid, and {bounds=}."
        )
    # the coalesced order: one block-diagonal import sorts every element by
    # (input, output) (rows grouped stably by input, storage order = output)
    gi = ii.to_host()[:T].reshape(D, 2 * m) + (np.arange(D) * n)[:, None]
    go = io.to_host()[:T].reshape(D, 2 * m) + (np.arange(D) * m)[:, None]
    from ._regrid import import_weights

    handle = import_weights(gi.ravel(), go.ravel(), w.to_host()[:T], D * n, D * m)
    ei, eo, ew = handle.export()
    parts = []
    for d in range(D):
        s, e = 2 * m * d, 2 * m * (d + 1)
        parts.append((ei[s:e] - d * n, eo[s:e] - d * m, ew[s:e].copy()))
    return _register_parts(parts, handle, n, m, shape_orthogonal)


def conservative_weights(coordinates_input, coordinates_output, axis_input, axis_output,
                         weights_input, perturb, seed):
    """The 1-D branch of ``_weights_conservative`` (``:11-102``)."""
    perturb = _util.resolve_perturb(perturb, coordinates_input)
    (coordinates_input, coordinates_output, axis_input, axis_output,
     shape_input, shape_output, shape_orthogonal) = _util.normalize_coordinates(
        coordinates_input=coordinates_input, coordinates_output=coordinates_output,
        axis_input=axis_input, axis_output=axis_output, perturb=perturb, seed=seed)
    shape_values_input = list(shape_input)
    for ax in axis_input:
        shape_values_input[ax] -= 1
    shape_values_input = tuple(shape_values_input)
    shape_values_output = list(shape_output)
    for ax in axis_output:
        shape_values_output[ax] -= 1
    shape_values_output = tuple(shape_values_output)
    if weights_input is not None:
        weights_input = np.broadcast_to(weights_input, shape_values_input)
    (x_input,) = coordinates_input
    (x_output,) = coordinates_output
    x_input = np.moveaxis(x_input, axis_input, ~0)
    x_output = np.moveaxis(x_output, axis_output, ~0)
    x_input = x_input.reshape(-1, x_input.shape[~0])
    x_output = x_output.reshape(-1, x_output.shape[~0])
    if weights_input is not None:
        weights_input = np.moveaxis(weights_input, axis_input, ~0)
        weights_input = weights_input.reshape(-1, weights_input.shape[~0])
    result = conservative_1d(x_input, x_output, weights_input, shape_orthogonal)
    return result, shape_values_input, shape_values_output


def multilinear_weights(coordinates_input, coordinates_output, axis_input, axis_output,
                        weights_input, bounds, perturb, seed):
    """``_weights_multilinear`` (``_weights_multilinear.py:9-145``)."""
    from ._find_indices import find_indices

    indices_output = find_indices(
        coordinates_input=coordinates_input,
        coordinates_output=coordinates_output,
        axis_input=axis_input,
        axis_output=axis_output,
        method="searchsorted",
    )
    if bounds not in ("extrapolate", "nan", "raise"):
        raise ValueError(
            f"Unrecognized {bounds=}, "
            f"expected one of ('extrapolate', 'nan', 'raise')."
        )
    if perturb is None:
        perturb = False
    (coordinates_input, coordinates_output, axis_input, axis_output,
     shape_input, shape_output, shape_orthogonal) = _util.normalize_coordinates(
        coordinates_input=coordinates_input, coordinates_output=coordinates_output,
        axis_input=axis_input, axis_output=axis_output, perturb=perturb, seed=seed)
    axis_input_last = ~np.arange(len(axis_input))[::-1]
    axis_output_last = ~np.arange(len(axis_output))[::-1]
    shape_input_last = tuple(shape_input[ax] for ax in axis_input)
    shape_output_last = tuple(shape_output[ax] for ax in axis_output)
    indices_output = tuple(
        np.moveaxis(v, axis_output, axis_output_last).reshape(-1, *shape_output_last)
        for v in indices_output
    )
    coordinates_input = tuple(
        np.moveaxis(v, axis_input, axis_input_last).reshape(-1, *shape_input_last)
        for v in coordinates_input
    )
    coordinates_output = tuple(
        np.moveaxis(v, axis_output, axis_output_last).reshape(-1, *shape_output_last)
        for v in coordinates_output
    )
    if weights_input is not None:
        weights_input = np.broadcast_to(weights_input, shape_input)
        weights_input = np.moveaxis(weights_input, axis_input, axis_input_last)
        weights_input = weights_input.reshape(-1, *shape_input_last)
    if len(axis_input) != 1:
        raise ValueError(
            f"{len(axis_input)}-dimensional multilinear interpolation is not supported"
        )
    (index,) = indices_output
    (x_input,) = coordinates_input
    (x_output,) = coordinates_output
    # find_indices broadcasts over the orthogonal axes of both grids
    index = np.broadcast_to(index, x_output.shape)
    result = multilinear_1d(index, x_input, x_output, weights_input, bounds, shape_orthogonal)
    return result, shape_input, shape_output
