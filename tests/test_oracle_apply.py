"""
Pins of the oracle's apply (``rgo_apply``, restating
``_regrid/_regrid_from_weights.py:165-182``): each output accumulates
``w[k] * x[in[k]]`` sequentially in the storage order of the triples,
starting from zero.

``numpy.add.at`` is an independent statement of exactly that: it is
unbuffered and applies its increments one index at a time, in order.  The
triples are shuffled, repeat (in, out) pairs and mix magnitudes so that a
sum in any other order (sorted, pairwise, by input) changes the bits; a
dense product pins the mathematics on a tiny case.
"""

import numpy as np
import pytest

from oracle import oracle


def _triples(rng, nnz, n_in, n_out):
    ii = rng.integers(0, n_in, nnz)
    io = rng.integers(0, n_out, nnz)
    # repeated pairs: the same (in, out) several times, not adjacent
    rep = rng.integers(0, nnz, nnz // 4)
    ii[rep[1:]] = ii[rep[0]]
    io[rep[1:]] = io[rep[0]]
    w = rng.uniform(-1, 1, nnz) * 10.0 ** rng.integers(-8, 9, nnz)
    perm = rng.permutation(nnz)
    return ii[perm], io[perm], w[perm]


@pytest.mark.parametrize("seed", range(6))
def test_apply_equals_add_at(seed):
    rng = np.random.default_rng(seed)
    n_in, n_out, D = 97, 61, 3
    ii, io, w = _triples(rng, 4000, n_in, n_out)
    x = rng.normal(size=(D, n_in)) * 10.0 ** rng.integers(-4, 5, (D, n_in))
    got = oracle.apply(ii, io, w, x, n_out)
    for d in range(D):
        ref = np.zeros(n_out)
        np.add.at(ref, io, w * x[d, ii])
        assert np.array_equal(got[d], ref)


def test_apply_is_order_sensitive():
    """The case is chosen so a reordered sum fails: the sorted order of the
    same triples gives other bits, so the pin above checks the order."""
    rng = np.random.default_rng(11)
    ii, io, w = _triples(rng, 4000, 97, 61)
    x = rng.normal(size=(1, 97)) * 10.0 ** rng.integers(-4, 5, (1, 97))
    a = oracle.apply(ii, io, w, x, 61)
    order = np.lexsort((io, ii))
    b = oracle.apply(ii[order], io[order], w[order], x, 61)
    assert not np.array_equal(a, b)


def test_apply_negative_indices_wrap():
    """numba indexing wraps negative indices (the 1-D builders emit them for
    descending grids): index -k reads element n - k."""
    rng = np.random.default_rng(3)
    n_in, n_out = 13, 7
    ii, io, w = _triples(rng, 200, n_in, n_out)
    x = rng.normal(size=(2, n_in))
    neg_i = ii - n_in
    neg_o = io - n_out
    assert np.array_equal(oracle.apply(neg_i, neg_o, w, x, n_out), oracle.apply(ii, io, w, x, n_out))


def test_apply_dense_product():
    rng = np.random.default_rng(5)
    n_in, n_out = 9, 6
    ii, io, w = _triples(rng, 40, n_in, n_out)
    w = np.abs(w) % 1.0 + 0.1  # moderate magnitudes: a tolerance comparison
    dense = np.zeros((n_out, n_in))
    np.add.at(dense, (io, ii), w)
    x = rng.normal(size=(4, n_in))
    np.testing.assert_allclose(oracle.apply(ii, io, w, x, n_out), x @ dense.T, rtol=1e-13,
                               atol=1e-13)
    # the swapped roles (the adjoint) are the other product
    y = rng.normal(size=(4, n_out))
    np.testing.assert_allclose(oracle.apply(io, ii, w, y, n_in), y @ dense, rtol=1e-13,
                               atol=1e-13)


def test_apply_empty_and_zero_batch():
    x = np.ones((2, 5))
    y = oracle.apply(np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0), x, 3)
    assert y.shape == (2, 3) and not y.any()
