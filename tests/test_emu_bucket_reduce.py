"""The per-cell reduction of the build (bucket_reduce_sorted, shared by the
GPU kernels and the CPU emulation) against numpy's np.add.reduceat, the
operation of the reference's coalesce (_weights/_weights_arrays.py:59-72):
runs of equal output cell in a sorted bucket reduce to their first term plus
the pairwise sum of the others, bit for bit.  Run lengths span the
sequential (< 8 further terms), eight-accumulator (8..128) and long (> 128,
left to the long-run pass) regimes, at the strides the kernels use."""
import numpy as np
import pytest

from tests.emu import emu


def _bucket(rng, lengths):
    outs = np.sort(rng.choice(10**6, size=len(lengths), replace=False)).astype(np.uint64)
    keys, vals, starts = [], [], []
    ordinal = 0
    for o, n in zip(outs, lengths):
        starts.append(len(keys))
        for _ in range(n):
            keys.append((int(o) << 32) | ordinal)
            ordinal += 1
        vals.extend(rng.normal(size=n) * 10.0 ** rng.integers(-8, 8, size=n))
    return np.array(keys, dtype=np.uint64), np.array(vals), np.array(starts), outs


@pytest.mark.parametrize("stride", [1, 64])
@pytest.mark.parametrize("seed", range(6))
def test_reduce_matches_reduceat(seed, stride):
    rng = np.random.default_rng(seed)
    lengths = rng.choice([1, 2, 3, 5, 7, 8, 9, 15, 16, 17, 31, 64, 100, 128, 129],
                         size=int(rng.integers(1, 40)))
    keys, vals, starts, outs = _bucket(rng, lengths)
    m, lr, k2, v2 = emu.bucket_reduce(keys, vals, stride)
    assert not lr
    assert m == len(lengths)
    assert np.array_equal(k2[:m], outs)
    assert np.array_equal(v2[:m], np.add.reduceat(vals, starts))


@pytest.mark.parametrize("stride", [1, 64])
def test_long_run_left_sorted(stride):
    rng = np.random.default_rng(9)
    lengths = [3, 130, 2]  # 130 terms: 129 after the first, past the 128 limit
    keys, vals, starts, outs = _bucket(rng, lengths)
    m, lr, k2, v2 = emu.bucket_reduce(keys, vals, stride)
    assert lr and m == 3
    assert np.array_equal(k2, keys) and np.array_equal(v2, vals)


def test_empty_and_single():
    m, lr, _, _ = emu.bucket_reduce(np.zeros(0, np.uint64), np.zeros(0))
    assert m == 0 and not lr
    m, lr, k2, v2 = emu.bucket_reduce(np.array([(7 << 32) | 3], np.uint64), np.array([-0.0]))
    assert m == 1 and k2[0] == 7 and np.signbit(v2[0])
