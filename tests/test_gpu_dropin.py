"""
The drop-in API on the GPU, tested the way the reference tests its public
functions (_weights_conservative_2d_test.py, _weights_test.py,
test_regrid_from_weights.py, _weights_transposed_test.py,
test_find_indices.py), plus bit-for-bit comparison with the reference
pipeline run on the oracle (tests/oracle_pipeline.py).
"""

import pickle

import numpy as np
import pytest

import regridding_amd as rg
from tests import oracle_pipeline as P
from tests.test_oracle_weights import KNOWN

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", range(len(KNOWN)))
def test_reference_known_answers(case):
    ci, co, values, wi, expected = KNOWN[case]
    co = (co[0] + 1e-6, co[1] + 1e-6)
    w = rg.weights(ci, co, weights_input=wi, method="conservative", perturb=False)
    result = rg.regrid_from_weights(*w, values_input=values)
    assert np.allclose(result, expected, rtol=1e-3)
    # and the same bits as the reference pipeline on the oracle
    wo = P.weights(ci, co, weights_input=wi, perturb=False)
    for a, b in zip(w[0][()], wo[0][()]):
        assert np.array_equal(a, b)
    assert np.array_equal(result, P.regrid_from_weights(*wo, values_input=values))


xb, yb = np.meshgrid(np.linspace(-1, 1, 11), np.linspace(-1, 1, 12), indexing="ij")
xo = xb * np.cos(0.2) - yb * np.sin(0.2)
yo = xb * np.sin(0.2) + yb * np.cos(0.2)
values_input = np.random.default_rng(42).random((10, 11))


@pytest.fixture(scope="module")
def weights():
    return rg.weights((xb, yb), (xo, yo), method="conservative")


def test_format(weights):
    for ii, io, v in weights[0].reshape(-1):
        assert ii.ndim == 1 and ii.shape == io.shape == v.shape
        assert np.issubdtype(ii.dtype, np.integer) and np.issubdtype(io.dtype, np.integer)
        assert np.issubdtype(v.dtype, np.floating)


def test_pairs_unique(weights):
    for ii, io, v in weights[0].reshape(-1):
        key = ii * (io.max() + 1) + io
        assert np.unique(key).size == key.size


def test_matches_oracle_pipeline(weights):
    wo = P.weights((xb, yb), (xo, yo))
    for a, b in zip(weights[0][()], wo[0][()]):
        assert np.array_equal(a, b)
    assert weights[1] == wo[1] and weights[2] == wo[2]
    assert np.array_equal(rg.regrid_from_weights(*weights, values_input=values_input),
                          P.regrid_from_weights(*wo, values_input=values_input))


def test_pickle(weights):
    unpickled = pickle.loads(pickle.dumps(weights))
    a = rg.regrid_from_weights(*unpickled, values_input=values_input)
    b = rg.regrid_from_weights(*weights, values_input=values_input)
    assert np.array_equal(a, b)


def test_regrid_from_weights_equals_regrid(weights):
    a = rg.regrid_from_weights(*weights, values_input=values_input)
    b = rg.regrid((xb, yb), (xo, yo), values_input, method="conservative")
    assert a.shape == b.shape
    assert np.array_equal(a, b)


def test_transpose_weights(weights):
    t = rg.transpose_weights(weights)
    assert t[1] == weights[2] and t[2] == weights[1]
    for tr, orig in zip(t[0].reshape(-1), weights[0].reshape(-1)):
        assert tr[0] is orig[1] and tr[1] is orig[0] and tr[2] is orig[2]
    # the device adjoint gives the reference's scatter over the swapped triples
    y = rg.regrid_from_weights(*weights, values_input=values_input)
    back = rg.regrid_from_weights(*t, values_input=y)
    ii, io, v = weights[0][()]
    from oracle import oracle
    expect = oracle.apply(io, ii, v, y.reshape(1, -1), values_input.size).reshape(values_input.shape)
    assert np.array_equal(back, expect)


def test_transpose_weights_conservative_round_trip(weights):
    wt = rg.transpose_weights_conservative(weights, (xb, yb), (xo, yo))
    assert wt[1] == weights[2] and wt[2] == weights[1]
    out = rg.regrid_from_weights(*weights, values_input=values_input)
    back = rg.regrid_from_weights(*wt, values_input=out)
    assert back.shape == values_input.shape
    assert back.sum() <= values_input.sum() * (1 + 1e-6)
    assert back.sum() >= 0.8 * values_input.sum()


def test_transpose_conservative_equals_reverse_build():
    x1 = np.linspace(-1, 1, num=10)[..., np.newaxis]
    y1 = np.linspace(-1, 1, num=11)
    x2 = np.linspace(-1, 1, num=5)[..., np.newaxis]
    y2 = np.linspace(-1, 1, num=6)
    w = rg.weights((x1, y1), (x2, y2), method="conservative")
    wt = rg.transpose_weights_conservative(w, (x1, y1), (x2, y2))
    expected = rg.weights((x2, y2), (x1, y1), method="conservative")
    a = rg.regrid_from_weights(*wt, values_input=np.array(10))
    b = rg.regrid_from_weights(*expected, values_input=np.array(10))
    assert np.allclose(a, b)


def test_transpose_conservative_inverts_weights_input_2d():
    rng = np.random.default_rng(0)
    gx, gy = np.linspace(-1, 1, num=8), np.linspace(-1, 1, num=12)
    x_in, y_in = np.meshgrid(gx, gy, indexing="ij")
    x_out, y_out = x_in + 0.03, y_in + 0.02
    ix, iy = np.arange(gx.size - 1), np.arange(gy.size - 1)
    wi = 1.0 + 0.2 * ix[:, None] + 0.05 * iy[None, :]
    values = rng.uniform(1, 2, size=(gx.size - 1, gy.size - 1))
    w = rg.weights((x_in, y_in), (x_out, y_out), weights_input=wi, method="conservative", perturb=False)
    wt = rg.transpose_weights_conservative(w, (x_in, y_in), (x_out, y_out), weights_input=wi)
    result = rg.regrid_from_weights(*wt, values_input=rg.regrid_from_weights(*w, values_input=values))
    wg = rg.weights((x_in, y_in), (x_out, y_out), method="conservative", perturb=False)
    wgt = rg.transpose_weights_conservative(wg, (x_in, y_in), (x_out, y_out))
    expected = rg.regrid_from_weights(
        *wgt, values_input=rg.regrid_from_weights(*wg, values_input=wi * values)) / wi
    assert np.allclose(result, expected)


def test_orthogonal_axis_elements():
    x, y = np.meshgrid(np.linspace(-1, 1, 10), np.linspace(-1, 1, 11), indexing="ij")
    ci = (x[..., None] + np.array([0, 0.001]), y[..., None] + np.array([0, 0.001]))
    co = (1.1 * (x[..., None] + np.array([0, 0.001])) + 0.01,
          1.2 * (y[..., None] + np.array([0, 0.01])) + 0.001)
    values = np.random.default_rng(1).normal(size=(9, 10, 2))
    w = rg.weights(ci, co, axis_input=(0, 1), axis_output=(0, 1), method="conservative")
    data = rg.regrid_from_weights(*w, values_input=values, axis_input=(0, 1), axis_output=(0, 1))
    wo = P.weights(ci, co, axis_input=(0, 1), axis_output=(0, 1))
    for k in range(2):
        for a, b in zip(w[0][k], wo[0][k]):
            assert np.array_equal(a, b)
    assert np.array_equal(data, P.regrid_from_weights(*wo, values_input=values,
                                                      axis_input=(0, 1), axis_output=(0, 1)))
    back = rg.regrid_from_weights(*rg.transpose_weights(w), values_input=data,
                                  axis_input=(0, 1), axis_output=(0, 1))
    assert back.shape == values.shape


def test_values_output_and_scalar_broadcast(weights):
    out = np.full((10, 11), 7.0)
    r = rg.regrid_from_weights(*weights, values_input=3.0, values_output=np.empty((10, 11)))
    assert r.shape == (10, 11)
    assert np.array_equal(r, rg.regrid_from_weights(*weights, values_input=np.full((10, 11), 3.0)))
    with pytest.raises(ValueError):
        rg.regrid_from_weights(*weights, values_input=values_input, values_output=out[:5])


def test_seed_semantics():
    g = ((xb, yb), (xo, yo))
    a = rg.weights(*g, method="conservative", seed=0)[0][()][2]
    b = rg.weights(*g, method="conservative", seed=1)[0][()][2]
    assert not np.array_equal(a, b)
    assert np.isclose(a.sum(), b.sum(), rtol=1e-6)
    c = rg.weights(*g, method="conservative", seed=None)[0][()][2]
    d = rg.weights(*g, method="conservative", seed=None)[0][()][2]
    assert not np.array_equal(c, d)


@pytest.mark.parametrize("method", ["brute", "searchsorted"])
@pytest.mark.parametrize("hi", [1, 2])
def test_find_indices_1d(method, hi):
    xi = np.linspace(-1, 1, num=32)
    xo = np.linspace(-hi, hi, num=64)
    (r,) = rg.find_indices((xi,), (xo,), method=method)
    where = (xo <= xi.max()) & (xo > xi.min())
    assert np.all(xi[r[where]] <= xo[where]) and np.all(xi[r[where] + 1] >= xo[where])


@pytest.mark.parametrize("method", ["brute", "searchsorted"])
def test_find_indices_outside_grid(method):
    x = np.array([0.0, 1.0, 2.0, 3.0, 4.0])
    (r,) = rg.find_indices((x,), (np.array([-1.0, 0.0, 0.5, 3.5, 4.0, 5.0]),), method=method)
    f = np.iinfo(int).max
    assert np.array_equal(r, np.array([f, 0, 0, 3, 3, f]))


def test_find_indices_fill_value():
    (r,) = rg.find_indices((np.array([0.0, 1.0, 2.0]),), (np.array([-1.0, 0.5, 7.0]),),
                           method="searchsorted", fill_value=-99)
    assert np.array_equal(r, np.array([-99, 0, -99]))


def test_find_indices_2d_secant():
    from oracle import oracle
    g = np.meshgrid(np.arange(3), np.arange(4), indexing="ij")
    pts = (np.array([[0.5, 1.5, 1.0], [-1.0, 1.2, 0.0]]), np.array([[0.5, 0.5, 4.5], [-1.0, 2.2, 0.0]]))
    i, j = rg.find_indices(g, pts, method="secant")
    for a in range(2):
        for b in range(3):
            e = oracle.index_of_point_secant((pts[0][a, b], pts[1][a, b]), g)
            assert (i[a, b], j[a, b]) == e


def test_returned_weights_are_writable_and_edits_apply():
    """The reference returns fresh writable arrays (_weights_arrays.py:68-72);
    an in-place edit of the values (or indices) must reach the next apply."""
    w = rg.weights((xb, yb), (xo, yo), method="conservative")
    ii, io, v = w[0][()]
    assert ii.flags.writeable and io.flags.writeable and v.flags.writeable
    before = rg.regrid_from_weights(*w, values_input=values_input)
    v *= 2.0
    after = rg.regrid_from_weights(*w, values_input=values_input)
    wo = P.weights((xb, yb), (xo, yo))
    oi, oo, ov = wo[0][()]
    edited = np.empty((), dtype=object)
    edited[()] = (oi, oo, ov * 2.0)
    assert np.array_equal(after, P.regrid_from_weights(edited, wo[1], wo[2],
                                                       values_input=values_input))
    assert not np.array_equal(before, after)
    # an index edit: output cell of the first entry moved
    io[0] = (io[0] + 1) % int(np.prod(w[2]))
    edited[()] = (oi, io.copy(), ov * 2.0)
    got = rg.regrid_from_weights(*w, values_input=values_input)
    assert np.array_equal(got, P.regrid_from_weights(edited, wo[1], wo[2],
                                                     values_input=values_input))
    # the transposed swap of the edited arrays applies the edited weights too
    wt = rg.transpose_weights(w)
    y = np.random.default_rng(3).random(w[2])
    et = np.empty((), dtype=object)
    et[()] = (io.copy(), oi, ov * 2.0)
    assert np.array_equal(rg.regrid_from_weights(*wt, values_input=y),
                          P.regrid_from_weights(et, wo[2], wo[1], values_input=y))
