"""
The build's cell search reads the interleaved (x, y) vertex copy
(index_secant_xy) with the arithmetic of the reference's secant search on
the separate arrays (index_secant, pinned against the oracle elsewhere): on
the CPU (tests/emu) both return the same cell, the same "outside" and the
same "exhaustive search" answer for points inside, outside and on the
vertices of warped grids.
"""

import ctypes

import numpy as np
import pytest

from regridding_amd import workloads
from tests.emu import emu

_pd = ctypes.POINTER(ctypes.c_double)
_pi = ctypes.POINTER(ctypes.c_int64)


def _pair(x, y, px, py):
    f = emu.lib().rge_secant_pair
    f.restype = None
    f.argtypes = [_pd, _pd, ctypes.c_int64, ctypes.c_int64, _pd, _pd, ctypes.c_int64, _pi, _pi]
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    px = np.ascontiguousarray(px, dtype=np.float64)
    py = np.ascontiguousarray(py, dtype=np.float64)
    a = np.empty(px.size, dtype=np.int64)
    b = np.empty(px.size, dtype=np.int64)
    f(x.ctypes.data_as(_pd), y.ctypes.data_as(_pd), x.shape[0], x.shape[1],
      px.ctypes.data_as(_pd), py.ctypes.data_as(_pd), px.size,
      a.ctypes.data_as(_pi), b.ctypes.data_as(_pi))
    return a, b


@pytest.mark.parametrize("shape,theta,warp", [((33, 29), 0.3, 0.05), ((70, 45), 0.7, 0.08),
                                              ((5, 5), 0.0, 0.0), ((2, 9), 1.2, 0.02)])
def test_interleaved_search_equals_separate(shape, theta, warp):
    x, y = workloads.warped_grid(shape, theta, warp, warp / 2)
    rng = np.random.default_rng(1)
    px = np.concatenate([rng.uniform(-1.6, 1.6, 400), x.ravel()[::3]])
    py = np.concatenate([rng.uniform(-1.6, 1.6, 400), y.ravel()[::3]])
    a, b = _pair(x, y, px, py)
    assert np.array_equal(a, b)
    assert (a >= 0).any() and (a < 0).any()
