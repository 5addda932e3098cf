"""
``fill()``: filling missing values by red-black Gauss-Seidel relaxation
(``regridding/_fill/_fill.py:10-109``, ``_fill/_gauss_seidel.py:12-80``).

The host part (broadcasting, the median starting guess, moving the relaxed
axes last) follows the reference; the sweeps run on the GPU
(``rg_fill_gauss_seidel_2d``) in the reference's update order.
"""

from __future__ import annotations

import warnings
from typing import Literal, Sequence

import numpy as np

from . import _util
from ._lib import DeviceArray, check, lib

__all__ = ["fill"]


def _guess_median(a: np.ndarray, where: np.ndarray, axis: tuple[int, ...]) -> np.ndarray:
    """The median of the valid elements of `a` along `axis`; slices with no
    valid element fall back to zero (``_gauss_seidel.py:58-80``)."""
    a = np.where(where, np.nan, a)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", category=RuntimeWarning)
        result = np.nanmedian(a, axis=axis, keepdims=True)
    return np.where(np.isnan(result), 0, result)


def fill_gauss_seidel(a: np.ndarray, where: np.ndarray, axis, guess=None,
                      num_iterations: int = 100) -> np.ndarray:
    a = a.copy()
    a, where = np.broadcast_arrays(a, where, subok=True)
    axis = _util.normalize_axis(axis, ndim=a.ndim)
    if guess is None:
        guess = _guess_median(a=a, where=where, axis=axis)
    a = np.array(a)  # broadcast views are read-only
    a[where] = np.broadcast_to(guess, a.shape)[where]
    if len(axis) != 2:
        raise ValueError(f"The number of interpolation axes, {len(axis)},is not supported")
    axis_last = ~np.arange(len(axis))[::-1]
    shape_last = tuple(a.shape[ax] for ax in axis)
    a = np.moveaxis(a, axis, axis_last)
    where = np.moveaxis(where, axis, axis_last)
    shape_moved = a.shape
    a = np.ascontiguousarray(a.reshape(-1, *shape_last), dtype=np.float64)
    w = np.ascontiguousarray(where.reshape(-1, *shape_last), dtype=np.uint8)
    T, ny, nx = a.shape
    if T and num_iterations > 0 and (nx == 1 or ny == 1):
        # the reference's grid spacing 2 / (n - 1) divides by zero here
        raise ZeroDivisionError("division by zero")
    da, dw = DeviceArray.from_host(a), DeviceArray.from_host(w)
    check(lib().rg_fill_gauss_seidel_2d(da.ptr, dw.ptr, T, ny, nx, int(num_iterations), 0))
    result = da.to_host().reshape(shape_moved)
    return np.moveaxis(result, axis_last, axis)


def fill(
    a: np.ndarray,
    where: None | np.ndarray = None,
    axis: None | int | Sequence[int] = None,
    method: Literal["gauss_seidel"] = "gauss_seidel",
    **kwargs,
) -> np.ndarray:
    """
    Fill an array with missing values by interpolating from the valid points.

    Same signature and result as ``regridding.fill``: ``where`` (default:
    the NaN elements) marks the missing values, relaxed over ``axis`` (two
    axes) with red-black Gauss-Seidel; keyword arguments ``guess`` (default:
    the median of the valid elements) and ``num_iterations`` (default 100).
    """
    if where is None:
        where = np.isnan(a)
    if method == "gauss_seidel":
        return fill_gauss_seidel(a=a, where=where, axis=axis, **kwargs)
    raise ValueError("Unrecognized method '{method}'")
