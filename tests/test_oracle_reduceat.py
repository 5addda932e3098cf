"""
Pins the coalescing reduction order: the oracle's (and the GPU's, through the
emulation library) per-segment sum must equal numpy.add.reduceat bit for bit
(_weights_arrays.py:72), for every segment length class of numpy's pairwise
summation (< 8, 8..128, > 128 with recursive splits).
"""

import numpy as np
import pytest

from oracle import oracle
from tests.emu import emu


def _hard_values(rng, n):
    # wide dynamic range and cancellation, so any other order shows up
    mant = rng.normal(size=n)
    expo = rng.integers(-30, 30, n)
    return mant * np.exp2(expo.astype(float))


LENGTHS = list(range(1, 40)) + [63, 64, 65, 127, 128, 129, 130, 135, 136, 137, 255, 256, 257,
                                 300, 1000, 1031, 4099]


@pytest.mark.parametrize("n", LENGTHS)
def test_reduce_segment_matches_numpy(n):
    rng = np.random.default_rng(n)
    for _ in range(20):
        a = _hard_values(rng, n)
        expected = np.add.reduceat(a, [0])[0]
        assert oracle.reduce_segment(a) == expected
        assert emu.reduceat_segment(a) == expected
        assert emu.reduceat_segment_inplace(a, 2) == expected


@pytest.mark.parametrize("n", [129, 130, 257, 1000, 4099, 20000, 65537, 300001])
def test_stackless_pairwise_matches_numpy(n):
    """The in-place post-order walk of the split tree (no stack) that the
    GPU's long-bucket path uses, against numpy, bit for bit."""
    rng = np.random.default_rng(n + 1)
    for stride in (1, 2):
        a = _hard_values(rng, n)
        assert emu.reduceat_segment_inplace(a, stride) == np.add.reduceat(a, [0])[0]


def test_reduceat_order_is_not_sequential():
    # a0 + pairwise(a1..): 1 + ((1e16 + -1e16) + 0) = 1, left-to-right gives 0
    a = np.array([1.0, 1e16, -1e16, 0.0])
    assert np.add.reduceat(a, [0])[0] == 1.0
    assert oracle.reduce_segment(a) == 1.0
    assert emu.reduceat_segment(a) == 1.0


def test_signed_zero():
    for a in ([-0.0], [-0.0, -0.0], [-0.0] * 9, [0.0, -0.0], [-0.0, 0.0]):
        a = np.array(a)
        expected = np.add.reduceat(a, [0])[0]
        got = oracle.reduce_segment(a)
        assert got == expected and np.signbit(got) == np.signbit(expected)
        got = emu.reduceat_segment(a)
        assert np.signbit(got) == np.signbit(expected)


def test_coalesce_matches_numpy_pipeline():
    # the oracle's coalesce against the reference's numpy statement of it
    rng = np.random.default_rng(3)
    n = 5000
    ii = rng.integers(3, 40, n)
    io = rng.integers(7, 30, n)
    w = _hard_values(rng, n)
    base_i, base_o = ii.min(), io.min()
    span = io.max() - base_o + 1
    key = (ii - base_i) * span + (io - base_o)
    order = np.argsort(key, kind="stable")
    key, vals = key[order], w[order]
    starts = np.flatnonzero(np.r_[True, key[1:] != key[:-1]])
    expected = (key[starts] // span + base_i, key[starts] % span + base_o,
                np.add.reduceat(vals, starts))
    got = oracle.coalesce(ii, io, w)
    for a, b in zip(got, expected):
        assert np.array_equal(a, b)
