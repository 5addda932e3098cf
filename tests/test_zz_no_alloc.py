"""
rg_find_cells_2d_noalloc, the cell search that allocates nothing (its
scratch is a fixed device array), against rg_find_cells_2d through the C
ABI: the same cells bit for bit on two calls in a row, the second smaller
(the array keeps the first call's contents), and a size beyond the fixed
maximum rejected before any launch.
"""

import ctypes

import numpy as np
import pytest

from regridding_amd import workloads
from regridding_amd._lib import DeviceArray, RGError, check, lib

pytestmark = pytest.mark.gpu

MAX_POINTS = 1 << 21  # RG_FIND_CELLS_NOALLOC_MAX


def _points(grid, n, seed):
    """Points over and around the grid: random ones (some outside), every
    vertex the draw reaches, and cell centres."""
    x, y = grid
    rng = np.random.default_rng(seed)
    px = rng.uniform(x.min() - 0.1, x.max() + 0.1, n)
    py = rng.uniform(y.min() - 0.1, y.max() + 0.1, n)
    k = min(n // 4, x.size)
    px[:k] = x.ravel()[:k]
    py[:k] = y.ravel()[:k]
    return px, py


def _search(fn, dx, dy, shape, px, py):
    n = px.size
    dpx, dpy = DeviceArray.from_host(px), DeviceArray.from_host(py)
    out = DeviceArray.empty((2 * n,), np.int64)
    check(fn(dx.ptr, dy.ptr, shape[0], shape[1], dpx.ptr, dpy.ptr, n, out.ptr, 0))
    return out.to_host()


def test_noalloc_equals_allocating_search_twice():
    L = lib()
    check(L.rg_init(0))
    grid = workloads.warped_grid((301, 257), 0.35, 0.06, 0.04)
    dx, dy = DeviceArray.from_host(grid[0]), DeviceArray.from_host(grid[1])
    for n, seed in ((MAX_POINTS, 1), (123_457, 2)):
        print(f"noalloc cell search: {n} points, seed {seed}", flush=True)
        px, py = _points(grid, n, seed)
        ref = _search(L.rg_find_cells_2d, dx, dy, grid[0].shape, px, py)
        got = _search(L.rg_find_cells_2d_noalloc, dx, dy, grid[0].shape, px, py)
        assert np.array_equal(got, ref)
        inside = ref[0::2] != np.iinfo(np.int64).max
        assert 0 < inside.sum() < n  # both inside and outside points


def test_noalloc_rejects_too_many_points():
    L = lib()
    grid = workloads.warped_grid(9, 0.1, 0.0, 0.0)
    dx, dy = DeviceArray.from_host(grid[0]), DeviceArray.from_host(grid[1])
    p = DeviceArray.empty((1,), np.float64)
    out = DeviceArray.empty((2,), np.int64)
    # the size check comes before any launch: no point array is read
    status = L.rg_find_cells_2d_noalloc(dx.ptr, dy.ptr, 9, 9, p.ptr, p.ptr, MAX_POINTS + 1,
                                        out.ptr, 0)
    assert status == 1  # RG_ERR_ARGUMENT
    assert b"exceed its fixed scratch" in L.rg_last_error()
    with pytest.raises(RGError):
        check(status)
