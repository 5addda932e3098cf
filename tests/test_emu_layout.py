"""
The graph workspace's layout (regridding_amd/csrc/rg_layout.hpp, run on the
CPU through tests/emu): live ranges never overlap, every range lies below
the reported peak, the same trace gives the same offsets (what lets a dry
run size the captured run's range), freed ranges are reused, and freeing
everything returns the layout to offset 0.
"""

import ctypes

import numpy as np
import pytest

from tests.emu import emu


def _trace(ops):
    L = emu.lib()
    f = L.rge_layout_trace
    f.restype = ctypes.c_int64
    f.argtypes = [ctypes.POINTER(ctypes.c_int64), ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]
    ops = np.ascontiguousarray(ops, dtype=np.int64)
    off = np.zeros(len(ops), dtype=np.int64)
    peak = f(ops.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), len(ops),
             off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    return off, peak


def _random_trace(rng, n):
    ops, live = [], []
    for _ in range(n):
        if live and rng.random() < 0.45:
            q = live.pop(int(rng.integers(len(live))))
            ops.append(-(q + 1))
        else:
            ops.append(256 * int(rng.integers(1, 64)) * (1 << int(rng.integers(0, 8))))
            live.append(len(ops) - 1)
    return ops, live


@pytest.mark.parametrize("seed", range(12))
def test_layout_random_traces(seed):
    rng = np.random.default_rng(seed)
    ops, _ = _random_trace(rng, 400)
    off, peak = _trace(ops)
    live = {}
    max_live = cur = 0
    for q, o in enumerate(ops):
        if o > 0:
            a, b = off[q], off[q] + o
            assert b <= peak
            for (c, d) in live.values():
                assert b <= c or d <= a, f"op {q} overlaps a live range"
            live[q] = (a, b)
            cur += o
            max_live = max(max_live, cur)
        else:
            a, b = live.pop(-o - 1)
            cur -= b - a
    assert max_live <= peak <= sum(o for o in ops if o > 0)
    off2, peak2 = _trace(ops)
    assert np.array_equal(off, off2) and peak == peak2


def test_layout_reuses_and_merges():
    # a, b, c; free a and b (merged); d fits in their place
    off, peak = _trace([1024, 2048, 512, -1, -2, 3072])
    assert list(off[:3]) == [0, 1024, 3072]
    assert off[5] == 0
    assert peak == 3584
    # a stream of same-sized temporaries costs one slot
    ops = []
    for k in range(50):
        ops += [4096, -(len(ops) + 1)]
    off, peak = _trace(ops)
    assert peak == 4096


def test_layout_free_all_returns_to_zero():
    rng = np.random.default_rng(3)
    ops, live = _random_trace(rng, 200)
    ops += [-(q + 1) for q in live] + [256]
    off, _ = _trace(ops)
    assert off[-1] == 0
