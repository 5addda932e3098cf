"""
Pins of the oracle pipeline's ``transpose_weights_conservative``
(``_weights_transposed.py:53-259``) on the CPU.

On unperturbed rectilinear grids the conservative weight of a cell pair is
``overlap / area_in``, so the renormalised transpose ``w * area_in /
area_out`` is ``overlap / area_out``: the weights of the reverse build
(``_weights_transposed_test.py:78-195``).  Non-uniform spacings make the
input and output areas differ cell by cell, so a swapped area ratio or a
dropped factor fails; the ``weights_input`` cases are the reference's
(``weights_input=2`` transposes to the reverse build with ``1/2``).
"""

import numpy as np
import pytest

from tests import oracle_pipeline as op


def _stretched(n, lo, hi, power):
    t = np.linspace(0, 1, n)
    return lo + (hi - lo) * t ** power


def _grids():
    xa = _stretched(9, -1.0, 1.0, 1.3)
    ya = _stretched(12, -1.0, 1.0, 0.8)
    xb = _stretched(6, -0.97, 1.02, 1.1)
    yb = _stretched(7, -1.03, 0.96, 1.2)
    ga = np.meshgrid(xa, ya, indexing="ij")
    gb = np.meshgrid(xb, yb, indexing="ij")
    return tuple(ga), tuple(gb)


def _dense(w, shape_in, shape_out):
    ii, io, v = w[0][()]
    d = np.zeros((int(np.prod(shape_out)), int(np.prod(shape_in))))
    np.add.at(d, (io, ii), v)
    return d


@pytest.mark.parametrize("wi", [None, 2.0, "field"])
def test_transpose_is_reverse_build(wi):
    ga, gb = _grids()
    shape_cells_a = (ga[0].shape[0] - 1, ga[0].shape[1] - 1)
    if isinstance(wi, str):
        ia, ja = np.meshgrid(np.arange(shape_cells_a[0]), np.arange(shape_cells_a[1]),
                             indexing="ij")
        wi = 1.0 + 0.3 * ia + 0.1 * ja
    w = op.weights(ga, gb, weights_input=wi, perturb=False)
    wt = op.transpose_weights_conservative(w, ga, gb, weights_input=wi)
    assert wt[1] == w[2] and wt[2] == w[1]
    rev = op.weights(gb, ga, perturb=False)
    got = _dense(wt, wt[1], wt[2])
    expect = _dense(rev, rev[1], rev[2])
    if wi is not None:
        # the transpose divides by the weighting once more: W(b->a) / wi
        expect = expect / np.broadcast_to(wi, shape_cells_a).reshape(-1, 1)
    np.testing.assert_allclose(got, expect, rtol=1e-12, atol=1e-15)


def test_swapped_area_ratio_fails():
    """The check above can tell the area ratio from its inverse."""
    ga, gb = _grids()
    w = op.weights(ga, gb, perturb=False)
    wt = op.transpose_weights_conservative(w, ga, gb)
    ii, io, v = w[0][()]
    io2, ii2, vt = wt[0][()]
    assert np.array_equal(io2, io) and np.array_equal(ii2, ii)
    ratio = vt / v
    # area_in / area_out varies from pair to pair (non-uniform cells)
    assert ratio.max() / ratio.min() > 1.5


def test_orthogonal_elements():
    """Elements along an orthogonal axis are renormalised with their own
    areas (the per-element volume slices of ``_weights_transposed.py:231-241``)."""
    ga, gb = _grids()
    ci = tuple(np.stack([g, 1.5 * g], axis=-1) for g in ga)
    co = tuple(np.stack([g, 1.5 * g], axis=-1) for g in gb)
    w = op.weights(ci, co, axis_input=(0, 1), axis_output=(0, 1), perturb=False)
    wt = op.transpose_weights_conservative(w, ci, co, axis_input=(0, 1), axis_output=(0, 1))
    for k in range(2):
        wk = op.weights(tuple(a[..., k] for a in ci), tuple(a[..., k] for a in co), perturb=False)
        wtk = op.transpose_weights_conservative(wk, tuple(a[..., k] for a in ci),
                                                tuple(a[..., k] for a in co))
        for a, b in zip(wt[0][k], wtk[0][()]):
            assert np.array_equal(a, b)
