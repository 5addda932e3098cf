"""
The oracle's 1-D weights (conservative sweep and multilinear) pinned by the
reference's own known answers (``_weights_conservative_1d_test.py:6-152``,
``_weights_multilinear_test.py:30-149``), by closed forms and by invariants.
"""

import numpy as np
import pytest

from oracle import oracle
from tests import oracle_pipeline as op

lin = np.linspace

C1D = [
    # x_input, x_output, values_input, values_expected, axis_input, axis_output, weights_input
    (lin(-1, 1, 11), lin(-1, 1, 6), np.ones(10), 2 * np.ones(5), None, None, None),
    (lin(-1, 1, 11), lin(-1, 1, 6) + 1e-6, np.ones(10), 2 * np.ones(5), None, None, None),
    (lin(-1, 1, 11), lin(-1, 1, 6) - 1e-6, np.ones(10), 2 * np.ones(5), None, None, None),
    (lin(1, -1, 11), lin(-1, 1, 21), np.ones(10), np.ones(20) / 2, 0, ~0, None),
    (lin(-1, 1, 11), lin(1, -1, 6), np.ones(10), 2 * np.ones(5), ~0, 0, None),
    (lin(1, -1, 11), lin(1, -1, 6), np.ones(10), 2 * np.ones(5), (0,), 0, None),
    (np.broadcast_to(lin(1, -1, 11), (3, 4, 11)), lin(1, -1, 6), np.ones(10), 2 * np.ones(5),
     ~0, 0, None),
    (np.broadcast_to(lin(1, -1, 11)[..., None, None], (11, 3, 4)), lin(1, -1, 6),
     np.ones(10)[..., None, None], 2 * np.ones(5), 0, ~0, None),
    (lin(1, -1, 11), np.broadcast_to(lin(1, -1, 6), (3, 4, 6)), np.ones(10), 2 * np.ones(5),
     ~0, ~0, None),
    (lin(-1, 1, 11), lin(1, 2, 6), np.ones(10), 0, None, None, None),
    (lin(1, -1, 11), lin(1, -1, 6), np.broadcast_to(np.ones(10), (3, 4, 10)), 2 * np.ones(5),
     ~0, ~0, None),
    (lin(-1, 1, 11), lin(-1, 1, 6), np.ones(10), 2 * np.ones(5), None, None, 1),
]


@pytest.mark.parametrize("case", range(len(C1D)))
def test_conservative_1d_known_answers(case):
    xi, xo, v, expected, ai, ao, wi = C1D[case]
    w = op.weights_conservative_1d(xi, xo, ai, ao, weights_input=wi)
    got = op.regrid_from_weights(*w, values_input=v, axis_input=ai, axis_output=ao)
    assert np.allclose(got, expected)


def test_conservative_1d_closed_form():
    # ascending grids: weight = overlap length / input cell length, computed
    # from the same two breakpoints as the sweep, so bit for bit
    rng = np.random.default_rng(3)
    for _ in range(20):
        xi = np.sort(rng.uniform(-1, 1, rng.integers(2, 30)))
        xo = np.sort(rng.uniform(-1.2, 1.2, rng.integers(2, 30)))
        if np.any(np.diff(xi) == 0) or np.any(np.diff(xo) == 0):
            continue
        ii, io, w = oracle.coalesce(*oracle.conservative_1d(xi, xo))
        exp = {}
        for a in range(xi.size - 1):
            for b in range(xo.size - 1):
                lo, hi = max(xi[a], xo[b]), min(xi[a + 1], xo[b + 1])
                if hi > lo:
                    exp[(a, b)] = (hi - lo) / (xi[a + 1] - xi[a])
        assert set(zip(ii.tolist(), io.tolist())) == set(exp)
        for p, q, v in zip(ii, io, w):
            assert v == exp[(int(p), int(q))]


def test_conservative_1d_conservation_and_reversal_indices():
    rng = np.random.default_rng(4)
    xi = np.sort(rng.uniform(-1, 1, 40))
    xo = np.linspace(-1.5, 1.5, 17)
    ii, io, w = oracle.coalesce(*oracle.conservative_1d(xi, xo))
    sums = np.zeros(xi.size - 1)
    np.add.at(sums, ii, w)
    assert np.allclose(sums, 1, rtol=0, atol=4e-16)  # input covered by the output
    # a descending input grid emits the reference's wraparound indices
    ri, ro, rw = oracle.coalesce(*oracle.conservative_1d(xi[::-1], xo))
    assert ri.min() < 0 and ri.max() < 0 and ro.min() >= 0
    wrapped = sorted(zip((ri + xi.size - 1).tolist(), ro.tolist()))
    assert wrapped == sorted(zip((xi.size - 2 - ii).tolist(), io.tolist()))


def test_conservative_1d_partial_and_disjoint():
    xi = np.linspace(0, 1, 11)
    assert oracle.conservative_1d(xi, np.linspace(2, 3, 5))[0].size == 0
    assert oracle.conservative_1d(xi, np.linspace(-3, -2, 5))[0].size == 0
    ii, io, w = oracle.coalesce(*oracle.conservative_1d(xi, np.linspace(0.55, 2, 4)))
    assert ii.min() == 5 and io.min() == 0


X = lin(-1, 1, 10)
Y = lin(-1, 1, 11)
XB, YB = np.meshgrid(X, Y, indexing="ij")
NY = lin(-1, 1, 5)
NX = lin(-1, 1, 6)
NXB, NYB = np.meshgrid(X, NY, indexing="ij")
NXB2, NYB2 = np.meshgrid(NX, Y, indexing="ij")

WML = [
    ((lin(-1, 1, 11),), (lin(-1, 1, 11),), np.square(lin(-1, 1, 11)), None, None, None,
     np.square(lin(-1, 1, 11))),
    ((Y,), (NY,), XB + YB, ~0, ~0, None, NXB + NYB),
    ((X[..., None],), (NX[..., None],), XB + YB, 0, 0, None, NXB2 + NYB2),
    ((X[..., None],), (0.1 * NX[..., None] + 0.001 * NY,), X[..., None], 0, 0, None,
     0.1 * NX[..., None] + 0.001 * NY),
    ((lin(-1, 1, 11),), (lin(-1, 1, 11),), np.square(lin(-1, 1, 11)), None, None, 1,
     np.square(lin(-1, 1, 11))),
    ((lin(-1, 1, 11),), (lin(-1, 1, 11),), np.square(lin(-1, 1, 11)), None, None, 0, 0),
]


@pytest.mark.parametrize("case", range(len(WML)))
def test_multilinear_known_answers(case):
    ci, co, v, ai, ao, wi, expected = WML[case]
    w = op.weights_multilinear(ci, co, ai, ao, weights_input=wi)
    got = op.regrid_from_weights(*w, values_input=v, axis_input=ai, axis_output=ao)
    assert np.allclose(got, expected)


def test_multilinear_linear_function_and_bounds():
    rng = np.random.default_rng(5)
    xi = np.sort(rng.uniform(-1, 1, 30))
    xo = rng.uniform(-1.5, 1.5, 50)
    idx = oracle.find_indices_1d(xi[None], xo[None], np.iinfo(int).max, "searchsorted")[0]
    ii, io, w, n_out = oracle.multilinear_1d(idx, xi, xo)
    y = oracle.apply(*oracle.coalesce(ii, io, w), (3 * xi - 1)[None], xo.size)[0]
    assert np.allclose(y, 3 * xo - 1, rtol=0, atol=1e-12)  # exact for linear data, extrapolated too
    assert n_out == np.sum((xo < xi[0]) | (xo >= xi[-1]))
    _, _, wn, _ = oracle.multilinear_1d(idx, xi, xo, bounds_nan=True)
    outside = np.repeat((xo < xi[0]) | (xo >= xi[-1]), 2)
    assert np.all(np.isnan(wn[outside])) and not np.any(np.isnan(wn[~outside]))
