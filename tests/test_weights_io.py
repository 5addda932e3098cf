"""
The weights file / wire format, the weights key and the persistent store,
on the host: exact round trips (values, dtypes, shapes, element order), the
documented layout, corrupt-file errors, key sensitivity, and the store's
hit / miss / eviction behaviour with the oracle pipeline as the builder.
"""

import os
import pickle

import numpy as np
import pytest

from regridding_amd import weights_io as wio
from tests import oracle_pipeline as op


def _random_weights(shape_orth=(2, 3), seed=0):
    r = np.random.default_rng(seed)
    arr = np.empty(shape_orth, dtype=object)
    for k in np.ndindex(*shape_orth):
        n = int(r.integers(0, 50)) if k != (0, 1) else 0
        arr[k] = (r.integers(-5, 100, n), r.integers(0, 70, n), r.normal(size=n))
    return arr, (4, 10, 10), (4, 7, 10)


def _assert_same(a, b):
    assert a[1] == b[1] and a[2] == b[2]
    assert a[0].shape == b[0].shape
    for x, y in zip(a[0].reshape(-1), b[0].reshape(-1)):
        for u, v in zip(x, y):
            assert np.asarray(v).dtype == (np.float64 if np.asarray(u).dtype.kind == "f" else np.int64)
            assert np.array_equal(np.asarray(u), np.asarray(v))


@pytest.mark.parametrize("mmap", [True, False])
def test_file_round_trip(tmp_path, mmap):
    w = _random_weights()
    p = tmp_path / "w.rgw"
    wio.save_weights(p, w)
    got = wio.load_weights(p, mmap=mmap)
    _assert_same(w, got)
    for element in got[0].reshape(-1):
        for a in element:
            assert not a.flags.writeable
    # the loaded tuple pickles like the reference's
    _assert_same(w, pickle.loads(pickle.dumps(got)))


def test_bytes_round_trip_and_layout():
    w = _random_weights((3,), seed=4)
    buf = wio.weights_to_bytes(w)
    _assert_same(w, wio.weights_from_bytes(buf))
    assert buf[:8] == b"RGWGHT01"
    L = int(np.frombuffer(buf[8:16], "<u8")[0])
    assert (16 + L) % 64 == 0
    import json

    h = json.loads(buf[16:16 + L])
    total = sum(h["nnz"])
    block = (8 * total + 63) // 64 * 64
    assert h["offsets"] == [0, block, 2 * block]
    assert len(buf) == 16 + L + 3 * block
    ii = np.frombuffer(buf, "<i8", total, 16 + L)
    assert np.array_equal(ii, np.concatenate([e[0] for e in w[0]]))
    wv = np.frombuffer(buf, "<f8", total, 16 + L + 2 * block)
    assert np.array_equal(wv, np.concatenate([e[2] for e in w[0]]))
    assert h["cells"] is None  # hand-made arrays: no device layout recorded


def test_empty_and_scalar_orthogonal_shape():
    arr = np.empty((), dtype=object)
    arr[()] = (np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0))
    w = (arr, (3, 3), (2, 2))
    _assert_same(w, wio.weights_from_bytes(wio.weights_to_bytes(w)))
    arr0 = np.empty((0, 2), dtype=object)
    w0 = (arr0, (0, 2, 3), (0, 2, 2))
    got = wio.weights_from_bytes(wio.weights_to_bytes(w0))
    assert got[0].shape == (0, 2)


def test_corrupt_files_raise(tmp_path):
    buf = wio.weights_to_bytes(_random_weights())
    with pytest.raises(ValueError):
        wio.weights_from_bytes(b"NOTWGHTS" + buf[8:])
    with pytest.raises(ValueError):
        wio.weights_from_bytes(buf[: len(buf) - 200])
    with pytest.raises(ValueError):
        wio.weights_from_bytes(buf[:40])
    bad = buf.replace(b'"version":1', b'"version":9')
    with pytest.raises(ValueError):
        wio.weights_from_bytes(bad)
    p = tmp_path / "empty.rgw"
    p.write_bytes(b"")
    with pytest.raises(ValueError):
        wio.load_weights(p)
    with pytest.raises(ValueError):
        elem = (np.zeros(3, np.int64), np.zeros(2, np.int64), np.zeros(3))
        arr = np.empty(1, dtype=object)
        arr[0] = elem
        wio.weights_to_bytes((arr, (3,), (3,)))


def _grids(n=9, m=7):
    xi, yi = np.meshgrid(np.linspace(-1, 1, n), np.linspace(-1, 1, n), indexing="ij")
    xo, yo = np.meshgrid(np.linspace(-0.8, 0.9, m), np.linspace(-0.9, 0.8, m), indexing="ij")
    return (xi, yi), (xo, yo)


def test_weights_key_sensitivity():
    gi, go = _grids()
    base = dict(method="conservative")
    k0 = wio.weights_key(gi, go, **base)
    assert k0 == wio.weights_key((gi[0].copy(), gi[1].copy()), go, **base)
    assert len(k0) == 32
    xi2 = gi[0].copy()
    xi2[3, 4] = np.nextafter(xi2[3, 4], 2)
    variants = [
        wio.weights_key((xi2, gi[1]), go, **base),
        wio.weights_key(gi, go, method="conservative", seed=43),
        wio.weights_key(gi, go, method="conservative", perturb=False),
        wio.weights_key(gi, go, method="conservative", weights_input=np.ones((8, 8))),
        wio.weights_key(gi, go, method="conservative", weights_input=2 * np.ones((8, 8))),
        wio.weights_key(gi, go, method="conservative", axis_input=(0, 1), axis_output=(0, 1)),
        wio.weights_key(gi, go, method="multilinear"),
        wio.weights_key(gi, go, method="conservative", bounds="nan"),
        wio.weights_key(go, gi, **base),
    ]
    keys = [k0] + variants
    assert len(set(keys)) == len(keys)
    # random state outside the arguments: not cacheable
    assert wio.weights_key(gi, go, method="conservative", seed=None) is None
    assert wio.weights_key(gi, go, method="conservative",
                           seed=np.random.default_rng(1)) is None
    # no perturbation: the seed plays no part
    assert wio.weights_key(gi, go, method="conservative", perturb=False, seed=None) == \
        wio.weights_key(gi, go, method="conservative", perturb=False, seed=7)


def test_store_hit_miss_and_eviction(tmp_path, monkeypatch):
    calls = []

    def build(ci, co, **kw):
        calls.append(kw)
        kw = {k: kw[k] for k in ("axis_input", "axis_output", "weights_input", "perturb", "seed")}
        return op.weights(ci, co, **kw)

    monkeypatch.setattr(wio, "_build", build)
    store = wio.WeightsStore(tmp_path / "store", device=False)
    gi, go = _grids()
    a = store.weights(gi, go, method="conservative")
    b = store.weights(gi, go, method="conservative")
    assert len(calls) == 1 and store.hits == 1 and store.misses == 1
    _assert_same(a, b)
    v = np.random.default_rng(0).uniform(size=(8, 8))
    assert np.array_equal(op.regrid_from_weights(*a, values_input=v),
                          op.regrid_from_weights(*b, values_input=v))
    # uncacheable seed: always built, nothing stored
    store.weights(gi, go, method="conservative", seed=None)
    store.weights(gi, go, method="conservative", seed=None)
    assert len(calls) == 3
    assert len(os.listdir(tmp_path / "store")) == 1
    # eviction keeps the directory under max_bytes (the newest file stays)
    size = os.path.getsize(store.path(wio.weights_key(gi, go, method="conservative")))
    small = wio.WeightsStore(tmp_path / "small", max_bytes=int(1.5 * size), device=False)
    for s in range(4):
        small.weights(gi, go, method="conservative", seed=s)
    files = os.listdir(tmp_path / "small")
    assert len(files) == 1
    assert files[0] == wio.weights_key(gi, go, method="conservative", seed=3) + ".rgw"
