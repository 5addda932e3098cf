"""
``transpose_weights()`` and ``transpose_weights_conservative()``
(``regridding/_weights/_weights_transposed/_weights_transposed.py:13-340``).
"""

from __future__ import annotations

import numpy as np

from . import _util
from ._lib import DeviceArray, check, lib

__all__ = ["transpose_weights", "transpose_weights_conservative", "cell_area_2d"]


def transpose_weights(
    weights: tuple[np.ndarray, tuple[int, ...], tuple[int, ...]],
) -> tuple[np.ndarray, tuple[int, ...], tuple[int, ...]]:
    """
    Swap the index arrays of every element, ``(i, j, w) -> (j, i, w)``,
    without copying (``_weights_transposed.py:13-50``).  When the arrays are
    device-resident, applying the result runs the adjoint kernel on the
    resident CSR by input.
    """
    weights_array, shape_input, shape_output = weights
    flat = weights_array.reshape(-1)
    result = np.empty(flat.size, dtype=object)
    for d in range(flat.size):
        indices_input, indices_output, values = flat[d]
        result[d] = (indices_output, indices_input, values)
    return result.reshape(weights_array.shape), shape_output, shape_input


def cell_area_2d(x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """Cell areas of an ``(nx, ny)`` vertex grid on the GPU
    (``grid_volume``, ``_weights_conservative_2d/_grids.py:50-75``)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    nx, ny = x.shape
    out = DeviceArray.empty(((nx - 1) * (ny - 1),), np.float64)
    dx, dy = DeviceArray.from_host(x), DeviceArray.from_host(y)
    check(lib().rg_cell_area_2d(dx.ptr, dy.ptr, nx, ny, out.ptr, 0))
    return out.to_host().reshape(nx - 1, ny - 1)


def _cell_volume(grid: tuple[np.ndarray, ...], axis: tuple[int, ...]) -> np.ndarray:
    """Cell volume of every element (``_weights_transposed.py:262-302``)."""
    shape = grid[0].shape
    axis_last = ~np.arange(len(axis))[::-1]
    shape_last = tuple(shape[ax] for ax in axis)
    if len(grid) == 2:
        x, y = grid
        x = np.moveaxis(x, axis, axis_last).reshape((-1,) + shape_last)
        y = np.moveaxis(y, axis, axis_last).reshape((-1,) + shape_last)
        result = np.empty((x.shape[0], shape_last[0] - 1, shape_last[1] - 1))
        for t in range(x.shape[0]):
            result[t] = cell_area_2d(x[t], y[t])
        # as _weights_transposed.py:296: the leading axes stay flattened
        result = np.reshape(result, x.shape[:-2] + result.shape[-2:])
        return np.moveaxis(result, axis_last, axis)
    if len(grid) == 1:
        (x,) = grid
        x = np.ascontiguousarray(np.moveaxis(x, axis, axis_last).reshape((-1,) + shape_last),
                                 dtype=np.float64)
        D, n = x.shape
        out = DeviceArray.empty((D, n - 1), np.float64)
        check(lib().rg_cell_length_1d(DeviceArray.from_host(x).ptr, n, D, out.ptr, 0))
        result = out.to_host()
        return np.moveaxis(result, axis_last, axis)
    raise ValueError("Grids greater than 2D not supported.")


def transpose_weights_conservative(
    weights: tuple[np.ndarray, tuple[int, ...], tuple[int, ...]],
    coordinates_input: np.ndarray | tuple[np.ndarray, ...],
    coordinates_output: np.ndarray | tuple[np.ndarray, ...],
    axis_input: None | int | tuple[int, ...] = None,
    axis_output: None | int | tuple[int, ...] = None,
    weights_input: None | np.ndarray = None,
) -> tuple[np.ndarray, tuple[int, ...], tuple[int, ...]]:
    """
    Transpose and renormalise weights so the reverse transform conserves
    flux and inverts `weights_input`: each value becomes
    ``w / weights_input[i]**2 * area_in[i] / area_out[o]`` with the areas of
    the unperturbed grids (``_weights_transposed.py:53-259``).
    """
    weights_array, shape_input, shape_output = weights
    (
        coordinates_input,
        coordinates_output,
        axis_input,
        axis_output,
        _,
        _,
        shape_orthogonal,
    ) = _util.normalize_coordinates(
        coordinates_input=coordinates_input,
        coordinates_output=coordinates_output,
        axis_input=axis_input,
        axis_output=axis_output,
    )
    # the stored flat indices use ascending-axis order
    axis_input = tuple(sorted(axis_input))
    axis_output = tuple(sorted(axis_output))
    axis_last_input = ~np.arange(len(axis_input))[::-1]
    axis_last_output = ~np.arange(len(axis_output))[::-1]
    size_orthogonal = int(np.prod(shape_orthogonal))

    if weights_input is not None:
        weights_input = np.broadcast_to(weights_input, shape_input)
        weights_input = np.moveaxis(weights_input, axis_input, axis_last_input)
        weights_input = weights_input.reshape(size_orthogonal, -1)

    volume_input = _cell_volume(coordinates_input, axis_input)
    volume_input = np.moveaxis(volume_input, axis_input, axis_last_input)
    volume_input = volume_input.reshape(size_orthogonal, -1)

    volume_output = _cell_volume(coordinates_output, axis_output)
    volume_output = np.moveaxis(volume_output, axis_output, axis_last_output)
    volume_output = volume_output.reshape(size_orthogonal, -1)

    flat = weights_array.reshape(-1)
    result = np.empty(flat.size, dtype=object)
    for d in range(flat.size):
        indices_input, indices_output, values = flat[d]
        values = np.array(values, dtype=np.float64)
        if weights_input is not None:
            values = values / np.square(weights_input[d][indices_input])
        values = values * volume_input[d][indices_input] / volume_output[d][indices_output]
        result[d] = (indices_output, indices_input, values)
    return result.reshape(weights_array.shape), shape_output, shape_input
