"""
The row start's winding number skips runs of 64 boundary-polygon edges
whose vertex y range excludes the point's y (k_row_start).  On the CPU
(tests/emu), the skipping sum equals the full sum of the reference's
winding terms for points at, just above and just below every vertex y, at
NaN vertices, and at random points, on star-shaped and grid-boundary
polygons.
"""

import ctypes

import numpy as np
import pytest

from oracle import oracle
from regridding_amd import workloads
from tests.emu import emu

_pd = ctypes.POINTER(ctypes.c_double)


def _winding(bpx, bpy, px, py, skip):
    f = emu.lib().rge_winding
    f.restype = ctypes.c_double
    f.argtypes = [_pd, _pd, ctypes.c_int64, ctypes.c_double, ctypes.c_double, ctypes.c_int]
    bpx = np.ascontiguousarray(bpx, dtype=np.float64)
    bpy = np.ascontiguousarray(bpy, dtype=np.float64)
    return f(bpx.ctypes.data_as(_pd), bpy.ctypes.data_as(_pd), bpx.size, px, py, skip)


def _star(n, rng):
    th = np.sort(rng.uniform(0, 2 * np.pi, n))
    r = rng.uniform(0.3, 1.0, n)
    return r * np.cos(th), r * np.sin(th)


def _grid_boundary(seed):
    x, y = workloads.warped_grid((70, 45), 0.3 + 0.1 * seed, 0.05, 0.04)
    return oracle.grid_boundary(x, y)


@pytest.mark.parametrize("seed", range(4))
def test_skipping_sum_equals_full_sum(seed):
    rng = np.random.default_rng(seed)
    polys = [_star(300, rng), _star(64, rng), _star(65, rng), _grid_boundary(seed)]
    for bpx, bpy in polys:
        pts = []
        for vy in bpy[:: max(1, bpy.size // 50)]:
            for py in (vy, np.nextafter(vy, np.inf), np.nextafter(vy, -np.inf)):
                pts.append((rng.uniform(bpx.min(), bpx.max()), py))
                pts.append((bpx[rng.integers(bpx.size)], py))
        pts += [tuple(p) for p in rng.uniform(-1.2, 1.2, (100, 2))]
        for px, py in pts:
            a = _winding(bpx, bpy, px, py, 0)
            b = _winding(bpx, bpy, px, py, 1)
            assert a == b, (px, py, a, b)


def test_nan_vertices_add_nothing():
    rng = np.random.default_rng(9)
    bpx, bpy = _star(200, rng)
    bpx = bpx.copy()
    bpy = bpy.copy()
    bpy[70:140] = np.nan  # a whole run and parts of two others
    for px, py in rng.uniform(-1, 1, (200, 2)):
        assert _winding(bpx, bpy, px, py, 0) == _winding(bpx, bpy, px, py, 1)
