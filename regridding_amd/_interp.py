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


def _positions(axis, ndim: int) -> list[int]:
    """Axis arguments as ascending non-negative positions."""
    if axis is None:
        return list(range(ndim))
    return sorted(int(ax) % ndim for ax in np.lib.array_utils.normalize_axis_tuple(axis, ndim))


def ndarray_linear_interpolation(
    a: np.ndarray,
    indices: tuple[np.ndarray, ...],
    axis: None | int | tuple[int, ...] = None,
    axis_indices: None | int | tuple[int, ...] = None,
):
    """
    Interpolate ``a`` at the fractional ``indices`` along ``axis`` (linear
    for one axis, bilinear for two), clamping to the first and last cell.
    Same signature, broadcasting, result and exception types as
    ``regridding.ndarray_linear_interpolation``
    (``_interp_ndarray.py:11-165``):

    * the k interpolated axes of ``a`` pair with the k index arrays in
      ascending axis position (whatever order ``axis`` lists them in);
    * the point axes of the indices (``axis_indices``, all by default) hold
      the evaluation points; every other axis of either operand is
      orthogonal, and the two operands must have the same number of
      orthogonal axes, which broadcast against each other position by
      position (``ValueError`` otherwise);
    * the result has the shape of the broadcast indices.

    The host only moves axes: the interpolated axes of ``a`` and the point
    axes of the indices go last, the orthogonal shapes broadcast, and one
    launch evaluates every point of every orthogonal element.
    """
    a = np.asarray(a)
    indices = tuple(np.asarray(ind) for ind in indices)
    ndim_i = np.broadcast(*indices).ndim
    ax_a = _positions(axis, a.ndim)
    ax_i = _positions(axis_indices, ndim_i)
    k = len(ax_a)
    if len(indices) != k:
        raise ValueError(
            f"ndarray_linear_interpolation: {len(indices)} index arrays were given for "
            f"{k} interpolated axes"
        )
    n_orth = a.ndim - k
    if ndim_i - len(ax_i) != n_orth:
        raise ValueError(
            f"ndarray_linear_interpolation: `a` has {n_orth} orthogonal axes but the indices "
            f"have {ndim_i - len(ax_i)}"
        )
    last_a = tuple(range(-k, 0))
    last_i = tuple(range(-len(ax_i), 0))
    a_m = np.moveaxis(a, ax_a, last_a)
    p_m = [np.moveaxis(np.broadcast_to(ind, np.broadcast(*indices).shape), ax_i, last_i)
           for ind in indices]
    grid_shape = a_m.shape[n_orth:]
    pts_shape = p_m[0].shape[n_orth:]
    shape_orth = np.broadcast_shapes(a_m.shape[:n_orth], p_m[0].shape[:n_orth])
    if k not in (1, 2):
        raise NotImplementedError("interpolation along more than two axes")
    D = int(np.prod(shape_orth))
    m = int(np.prod(pts_shape))
    a_flat = np.ascontiguousarray(
        np.broadcast_to(a_m, shape_orth + grid_shape).reshape((D,) + grid_shape),
        dtype=np.float64)
    p_flat = [np.ascontiguousarray(np.broadcast_to(p, shape_orth + pts_shape).reshape(D, m),
                                   dtype=np.float64) for p in p_m]
    out = DeviceArray.empty((D, m), np.float64)
    da = DeviceArray.from_host(a_flat)
    dp = [DeviceArray.from_host(p) for p in p_flat]
    if k == 1:
        check(lib().rg_interp_linear_1d(da.ptr, grid_shape[0], dp[0].ptr, m, D, out.ptr, 0))
    else:
        check(lib().rg_interp_linear_2d(da.ptr, grid_shape[0], grid_shape[1], dp[0].ptr,
                                        dp[1].ptr, m, D, out.ptr, 0))
    res = out.to_host().reshape(shape_orth + pts_shape)
    return np.moveaxis(res, last_i, ax_i)
