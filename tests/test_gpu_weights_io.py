"""
Weights written by the GPU builders, saved and loaded back onto the device:
the loaded arrays apply bit-identically to the originals and are found in
the device registry (no second upload), for a 2-D conservative pair, 1-D
weights over orthogonal axes (one block-diagonal handle) and transposed
weights; the persistent store returns the same weights on a hit.
"""

import numpy as np
import pytest

import regridding_amd as rg
from regridding_amd import _cache
from regridding_amd import weights_io as wio

pytestmark = pytest.mark.gpu


def _grids(n=33, m=25):
    r = np.random.default_rng(n)
    xi, yi = np.meshgrid(np.linspace(-1, 1, n), np.linspace(-1, 1, n), indexing="ij")
    xi = xi + 0.02 * np.sin(3 * yi)
    xo, yo = np.meshgrid(np.linspace(-0.9, 0.9, m), np.linspace(-0.9, 0.9, m), indexing="ij")
    return (xi, yi), (xo, yo), r.uniform(size=(3, n - 1, n - 1))


def test_conservative_2d_save_load_device(tmp_path):
    gi, go, v = _grids()
    w = rg.weights(gi, go, method="conservative")
    p = tmp_path / "w.rgw"
    wio.save_weights(p, w)
    got = wio.load_weights(p, device=True)
    ii, io, vals = got[0][()]
    assert _cache.lookup(ii, io, vals) is not None
    assert np.array_equal(rg.regrid_from_weights(*got, values_input=v),
                          rg.regrid_from_weights(*w, values_input=v))
    # transposed weights keep the swapped cell counts
    wt = rg.transpose_weights(w)
    wio.save_weights(tmp_path / "t.rgw", wt)
    gt = wio.load_weights(tmp_path / "t.rgw", device=True)
    y = np.random.default_rng(1).uniform(size=(2, 24, 24))
    assert np.array_equal(rg.regrid_from_weights(*gt, values_input=y),
                          rg.regrid_from_weights(*wt, values_input=y))
    # bytes on the wire, device import
    gb = wio.weights_from_bytes(wio.weights_to_bytes(w), device=True)
    assert np.array_equal(rg.regrid_from_weights(*gb, values_input=v),
                          rg.regrid_from_weights(*w, values_input=v))


def test_conservative_1d_orthogonal_block_handle(tmp_path):
    r = np.random.default_rng(2)
    xi = np.sort(r.uniform(-1, 1, (4, 3, 41)), axis=-1)
    xo = np.linspace(-0.9, 0.9, 17)
    w = rg.weights(xi, xo, axis_input=-1, axis_output=-1, method="conservative")
    assert w[0].shape == (4, 3)
    wio.save_weights(tmp_path / "c1.rgw", w)
    got = wio.load_weights(tmp_path / "c1.rgw", device=True)
    v = r.uniform(size=(4, 3, 40))
    assert np.array_equal(rg.regrid_from_weights(*got, values_input=v, axis_input=-1,
                                                 axis_output=-1),
                          rg.regrid_from_weights(*w, values_input=v, axis_input=-1,
                                                 axis_output=-1))
    ii, io, vals = got[0][1, 2]
    hit = _cache.lookup_part(ii, io, vals)
    assert hit is not None and hit[2][:2] == (5, 12)


def test_store_device_hits(tmp_path):
    gi, go, v = _grids(41, 29)
    store = wio.WeightsStore(tmp_path / "s")
    a = store.weights(gi, go, method="conservative")
    b = store.weights(gi, go, method="conservative")
    assert store.hits == 1 and store.misses == 1
    for x, y in zip(a[0][()], b[0][()]):
        assert np.array_equal(x, y)
    vv = np.random.default_rng(3).uniform(size=(2, 40, 40))
    assert np.array_equal(rg.regrid_from_weights(*a, values_input=vv),
                          rg.regrid_from_weights(*b, values_input=vv))
    assert _cache.lookup(*b[0][()]) is not None
