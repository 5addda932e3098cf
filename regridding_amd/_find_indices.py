"""
``find_indices()``: the drop-in cell search
(``regridding/_find_indices/_find_indices.py:12-136``).

1-D ``"brute"`` and ``"searchsorted"`` run on the GPU with the reference's
semantics and sentinels.  For 2-D grids the reference's methods raise
``ValueError`` (``_find_indices_brute.py:17-19``,
``_find_indices_searchsorted.py:17-19``) and still do here; the extra method
``"secant"`` exposes the 2-D search the conservative build uses
(``index_of_point_secant``, ``_weights_conservative_2d/_grids.py:356-463``):
the lowest-index cell containing each output vertex.
"""

from __future__ import annotations

from typing import Literal, Sequence

import numpy as np

from . import _util
from ._lib import DeviceArray, check, lib, RG_FIND_BRUTE, RG_FIND_SEARCHSORTED

__all__ = ["find_indices"]


def _find_1d(x_in: np.ndarray, x_out: np.ndarray, fill_value: int, method: int) -> np.ndarray:
    a = np.ascontiguousarray(x_in, dtype=np.float64)
    v = np.ascontiguousarray(x_out, dtype=np.float64)
    D, m = a.shape
    _, n = v.shape
    da, dv = DeviceArray.from_host(a), DeviceArray.from_host(v)
    dr = DeviceArray.empty((D, n), np.int64)
    check(lib().rg_find_indices_1d(da.ptr, m, dv.ptr, n, D, method, int(fill_value), dr.ptr, 0))
    return dr.to_host()


def _find_2d(grid: tuple[np.ndarray, np.ndarray], points: tuple[np.ndarray, np.ndarray],
             fill_value: int) -> tuple[np.ndarray, np.ndarray]:
    """grid: (D, nx, ny) x2; points: (D, n) x2 -> (D, n) int64 x2."""
    D = grid[0].shape[0]
    n = points[0].shape[1]
    ri = np.empty((D, n), np.int64)
    rj = np.empty((D, n), np.int64)
    for d in range(D):
        x = np.ascontiguousarray(grid[0][d], dtype=np.float64)
        y = np.ascontiguousarray(grid[1][d], dtype=np.float64)
        px = np.ascontiguousarray(points[0][d], dtype=np.float64)
        py = np.ascontiguousarray(points[1][d], dtype=np.float64)
        dx, dy, dpx, dpy = (DeviceArray.from_host(a) for a in (x, y, px, py))
        dij = DeviceArray.empty((n, 2), np.int64)
        check(lib().rg_find_cells_2d(dx.ptr, dy.ptr, x.shape[0], x.shape[1],
                                     dpx.ptr, dpy.ptr, n, dij.ptr, 0))
        ij = dij.to_host()
        outside = ij[:, 0] == np.iinfo(np.int64).max
        ij[outside] = fill_value
        ri[d], rj[d] = ij[:, 0], ij[:, 1]
    return ri, rj


def find_indices(
    coordinates_input: tuple[np.ndarray, ...],
    coordinates_output: tuple[np.ndarray, ...],
    axis_input: None | int | Sequence[int] = None,
    axis_output: None | int | Sequence[int] = None,
    fill_value: None | int = None,
    method: Literal["brute", "searchsorted", "secant"] = "brute",
) -> tuple[np.ndarray, ...]:
    """
    Find the index of the input cell containing each output vertex.

    Same signature, result and errors as ``regridding.find_indices``; one
    int64 array per resampled axis with the output grid's shape, `fill_value`
    (default ``numpy.iinfo(int).max``) outside the input grid.
    """
    (
        coordinates_input,
        coordinates_output,
        axis_input_,
        axis_output_,
        shape_input,
        shape_output,
        shape_orthogonal,
    ) = _util.normalize_coordinates(
        coordinates_input=coordinates_input,
        coordinates_output=coordinates_output,
        axis_input=axis_input,
        axis_output=axis_output,
    )

    if fill_value is None:
        fill_value = np.iinfo(int).max

    if method == "secant" and len(coordinates_input) == 2:
        # the 2-D extension keeps the grid's own axis order (ascending), so
        # (i, j) index the input arrays as given
        axis_input_ = tuple(sorted(axis_input_))
        axis_output_ = tuple(sorted(axis_output_))

    axis_input_last = ~np.arange(len(axis_input_))[::-1]
    axis_output_last = ~np.arange(len(axis_output_))[::-1]
    shape_input_last = tuple(shape_input[ax] for ax in axis_input_)
    shape_output_last = tuple(shape_output[ax] for ax in axis_output_)

    coordinates_input = tuple(
        np.moveaxis(v, axis_input_, axis_input_last).reshape(-1, *shape_input_last)
        for v in coordinates_input
    )
    coordinates_output = tuple(
        np.moveaxis(v, axis_output_, axis_output_last).reshape(-1, *shape_output_last)
        for v in coordinates_output
    )

    ndim = len(coordinates_input)
    if method in ("brute", "searchsorted"):
        if ndim != 1:
            label = "brute-force" if method == "brute" else "searchsorted"
            raise ValueError(f"{ndim}-dimensional {label} search not supported")
        code = RG_FIND_BRUTE if method == "brute" else RG_FIND_SEARCHSORTED
        indices = (_find_1d(coordinates_input[0], coordinates_output[0], fill_value, code),)
    elif method == "secant":
        if ndim != 2:
            raise ValueError(f"{ndim}-dimensional secant search not supported")
        (xi, yi), (xo, yo) = coordinates_input, coordinates_output
        n = int(np.prod(xo.shape[1:]))
        indices = _find_2d((xi, yi), (xo.reshape(-1, n), yo.reshape(-1, n)), fill_value)
    else:
        raise ValueError(f"method `{method}` not recognized.")

    return tuple(
        np.moveaxis(
            a=i.reshape(*shape_orthogonal, *shape_output_last),
            source=axis_output_last,
            destination=axis_output_,
        )
        for i in indices
    )
