"""The register sorting network of the kernels (``sort_network`` in
``rg_sweep.hpp``: the group reduction sorts each cell's keys with it, the
row sort each output row's columns).

By the 0-1 principle a comparator network sorts every input if and only if
it sorts every sequence of zeros and ones, so the network is checked on all
2^16 (and 2^8) of them; random keys with their positions as values check
that every value travels with its key.
"""

import itertools

import numpy as np

from tests.emu import emu


def test_sorts_every_zero_one_sequence():
    for n in (8, 16):
        bits = np.array(list(itertools.product((0, 1), repeat=n)), dtype=np.uint64)
        k, _ = emu.sort_network(bits, n)
        assert np.array_equal(k, np.sort(bits, axis=1))


def test_values_follow_their_keys():
    rng = np.random.default_rng(3)
    for n in (8, 16):
        keys = np.stack([rng.choice(2**40, n, replace=False) for _ in range(2000)]).astype(np.uint64)
        k, pos = emu.sort_network(keys, n)
        assert np.array_equal(k, np.sort(keys, axis=1))
        assert np.array_equal(np.take_along_axis(keys, pos, axis=1), k)
    # padding keys (all ones) sort last
    keys = np.full((1, 16), np.iinfo(np.uint64).max, dtype=np.uint64)
    keys[0, :5] = [9, 3, 7, 1, 5]
    k, _ = emu.sort_network(keys, 16)
    assert list(k[0, :5]) == [1, 3, 5, 7, 9]
