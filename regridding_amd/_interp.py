"""
``ndarray_linear_interpolation()`` (``regridding/_interp_ndarray.py:11-165``):
an array evaluated at fractional indices along one or two of its axes.  The
broadcasting of the array against the indices follows the reference; the
evaluation runs on the GPU (``rg_interp_linear_1d`` / ``_2d``), one lane per
output point.
"""

from __future__ import annotations

import numpy as np

from ._lib import DeviceArray, check, lib

__all__ = ["ndarray_linear_interpolation"]


def _normalize_axis_tuple(axis, ndim):
    return np.lib.array_utils.normalize_axis_tuple(axis, ndim=ndim)


def ndarray_linear_interpolation(
    a: np.ndarray,
    indices: tuple[np.ndarray, ...],
    axis: None | int | tuple[int, ...] = None,
    axis_indices: None | int | tuple[int, ...] = None,
):
    """
    Interpolate ``a`` at the fractional ``indices`` along ``axis`` (linear
    for one axis, bilinear for two), clamping to the first and last cell.
    Same signature, broadcasting and errors as
    ``regridding.ndarray_linear_interpolation``.
    """
    a = np.asarray(a)
    shape_a = a.shape
    ndim_a = a.ndim
    indices_broadcasted = np.broadcast(*indices)
    shape_indices = indices_broadcasted.shape
    ndim_indices = indices_broadcasted.ndim

    if axis is None:
        axis = tuple(range(ndim_a))
    axis = _normalize_axis_tuple(axis, ndim_a)
    axis = tuple(~(~np.array(axis) % ndim_a))
    if axis_indices is None:
        axis_indices = tuple(range(ndim_indices))
    axis_indices = _normalize_axis_tuple(axis_indices, ndim_indices)
    axis_indices = tuple(~(~np.array(axis_indices) % ndim_indices))

    if len(indices) != len(axis):
        raise ValueError(
            f"The number of indices, {len(indices)}, must match the number of elements "
            f"in axis, {len(axis)}"
        )

    axis_orthogonal_a = tuple(ax for ax in range(-ndim_a, 0) if ax not in axis)
    axis_orthogonal_indices = tuple(ax for ax in range(-ndim_indices, 0) if ax not in axis_indices)
    shape_orthogonal_a = tuple(shape_a[ax] for ax in axis_orthogonal_a)
    shape_orthogonal_indices = tuple(shape_indices[ax] for ax in axis_orthogonal_indices)
    shape_orthogonal = np.broadcast_shapes(shape_orthogonal_a, shape_orthogonal_indices)

    ndim_broadcasted_a = len(shape_orthogonal) + len(axis)
    ndim_broadcasted_indices = len(shape_orthogonal) + len(axis_indices)
    shape_broadcasted_a = tuple(
        shape_a[ax] if ax in axis else shape_orthogonal[axis_orthogonal_a.index(ax)]
        for ax in range(-ndim_broadcasted_a, 0)
    )
    shape_broadcasted_indices = tuple(
        shape_indices[ax] if ax in axis_indices
        else shape_orthogonal[axis_orthogonal_indices.index(ax)]
        for ax in range(-ndim_broadcasted_indices, 0)
    )
    a = np.broadcast_to(a, shape_broadcasted_a)
    indices = tuple(np.broadcast_to(ind, shape_broadcasted_indices) for ind in indices)

    if len(axis) not in (1, 2):
        raise NotImplementedError

    # orthogonal axes first (in the reference's ndindex order), then the
    # interpolated axes of `a` in ascending order -- the order in which the
    # reference's a[index_a] presents them to the evaluator -- and the point
    # axes of the indices
    orth_a = tuple(ax for ax in range(-ndim_broadcasted_a, 0) if ax not in axis)
    orth_i = tuple(ax for ax in range(-ndim_broadcasted_indices, 0) if ax not in axis_indices)
    axis_a = tuple(ax for ax in range(-ndim_broadcasted_a, 0) if ax in axis)
    a_m = np.moveaxis(a, orth_a + axis_a, tuple(range(ndim_broadcasted_a)))
    D = int(np.prod(shape_orthogonal))
    a_m = np.ascontiguousarray(
        a_m.reshape((D,) + tuple(shape_broadcasted_a[ax] for ax in axis_a)), dtype=np.float64)
    pts_shape = tuple(shape_broadcasted_indices[ax] for ax in axis_indices)
    m = int(np.prod(pts_shape))
    ind_m = [
        np.ascontiguousarray(
            np.moveaxis(ind, orth_i + axis_indices, tuple(range(ndim_broadcasted_indices)))
            .reshape(D, m), dtype=np.float64)
        for ind in indices
    ]
    out = DeviceArray.empty((D, m), np.float64)
    da = DeviceArray.from_host(a_m)
    dx = [DeviceArray.from_host(v) for v in ind_m]
    if len(axis) == 1:
        check(lib().rg_interp_linear_1d(da.ptr, a_m.shape[1], dx[0].ptr, m, D, out.ptr, 0))
    else:
        check(lib().rg_interp_linear_2d(da.ptr, a_m.shape[1], a_m.shape[2], dx[0].ptr,
                                        dx[1].ptr, m, D, out.ptr, 0))
    res = out.to_host().reshape(tuple(shape_orthogonal) + pts_shape)
    return np.moveaxis(res, tuple(range(ndim_broadcasted_indices)), orth_i + axis_indices)
