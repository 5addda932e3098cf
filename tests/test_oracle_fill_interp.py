"""
The oracle's ``fill`` (red-black Gauss-Seidel) and linear interpolation,
pinned by the reference's own tests (``_fill/_fill_test.py:11-150``,
``_tests/test_interp_ndarray.py:10-97``: scipy.ndimage.map_coordinates), by
the fixed point of the relaxation (a sparse linear solve), by a vectorised
red-black restatement, and by closed forms (linear / bilinear functions are
reproduced exactly, inside and outside the grid).
"""

import numpy as np
import pytest
import scipy.ndimage
import scipy.sparse
import scipy.sparse.linalg

from oracle import oracle

NX, NY, NT = 11, 12, 13
rng = np.random.default_rng(7)


def _fill_cases():
    r = np.random.default_rng(11)
    with np.errstate(invalid="ignore"):
        sq = np.sqrt(r.uniform(-0.1, 1, size=(NX, NT, NY)))
    return [
        (r.uniform(0, 1, size=(NX, NY)), r.uniform(0, 1, size=(NX, NY)) > 0.9, None),
        (r.uniform(0, 1, size=(NT, NX, NY)), r.uniform(0, 1, size=(NT, NX, NY)) > 0.9, (~1, ~0)),
        (sq, None, (0, ~0)),
    ]


@pytest.mark.parametrize("case", range(3))
@pytest.mark.parametrize("guess", [None, 0.5])
def test_fill_reference_known_answers(case, guess):
    """_fill_test.py:11-57"""
    a, where, axis = _fill_cases()[case]
    result = oracle.fill(a, where, axis, guess=guess, num_iterations=11)
    if where is None:
        where = np.isnan(a)
    assert np.all(np.isfinite(result))
    assert np.allclose(result[~where], a[~where])
    assert np.all(result[where] != 0)


@pytest.mark.parametrize("guess", [None, 2.0, np.arange(NT).reshape(NT, 1, 1)])
def test_fill_guess_zero_iterations(guess):
    """_fill_test.py:60-90"""
    a = rng.uniform(0, 1, size=(NT, NX, NY))
    where = rng.uniform(0, 1, size=a.shape) > 0.9
    result = oracle.fill(a, where, (~1, ~0), guess=guess, num_iterations=0)
    if guess is None:
        guess = np.nanmedian(np.where(where, np.nan, a), axis=(~1, ~0), keepdims=True)
    assert np.allclose(result[~where], a[~where])
    assert np.allclose(result[where], np.broadcast_to(guess, a.shape)[where])


def test_fill_missing_cluster_harmonic():
    """_fill_test.py:93-109: a harmonic field is recovered."""
    x = np.linspace(-1, 1, num=32)
    a = x[:, None] + 2 * x[None, :]
    am = a.copy()
    am[10:16, 10:16] = np.nan
    result = oracle.fill(am, num_iterations=1000)
    assert np.all(np.isfinite(result))
    assert np.allclose(result, a, atol=1e-4)


def test_fill_all_missing_slice():
    """_fill_test.py:112-130"""
    a = rng.uniform(0, 1, size=(NT, NX, NY))
    where = rng.uniform(0, 1, size=a.shape) > 0.9
    where[0] = True
    result = oracle.fill(a, where, (~1, ~0), num_iterations=11)
    assert np.all(np.isfinite(result))
    assert np.all(result[0] == 0)
    assert np.allclose(result[~where], a[~where])


@pytest.mark.parametrize("shape", [(4, 6), (5, 7), (7, 4), (6, 9), (2, 3)])
def test_fill_fixed_point_is_the_discrete_laplace_solution(shape):
    """Many sweeps converge to the solution of the 5-point periodic Laplace
    equations on the missing cells (coefficients 1/dx^2, 1/dy^2 with
    dx = 2/(nx-1), dy = 2/(ny-1)), solved independently with a sparse LU."""
    ny, nx = shape
    r = np.random.default_rng(ny * 100 + nx)
    a = r.uniform(-1, 1, size=(1, ny, nx))
    where = r.uniform(size=a.shape) > 0.6
    where[0, 0, 0] = False
    cx, cy = 1 / (2 / (nx - 1)) ** 2, 1 / (2 / (ny - 1)) ** 2
    cells = [(j, i) for j in range(ny) for i in range(nx) if where[0, j, i]]
    pos = {c: k for k, c in enumerate(cells)}
    A = scipy.sparse.lil_matrix((len(cells), len(cells)))
    b = np.zeros(len(cells))
    for k, (j, i) in enumerate(cells):
        A[k, k] += 2 * (cx + cy)
        for (jj, ii), c in (((j, i - 1), cx), ((j, i + 1), cx), ((j - 1, i), cy), ((j + 1, i), cy)):
            jj, ii = jj % ny, ii % nx
            if where[0, jj, ii]:
                A[k, pos[(jj, ii)]] -= c
            else:
                b[k] += c * a[0, jj, ii]
    sol = scipy.sparse.linalg.spsolve(A.tocsc(), b)
    got = oracle.fill_gauss_seidel_2d(a, where, 4000)
    np.testing.assert_allclose(got[0][where[0]], sol, rtol=0, atol=1e-9)
    assert np.array_equal(got[0][~where[0]], a[0][~where[0]])


@pytest.mark.parametrize("shape", [(4, 6), (8, 2), (6, 6)])
def test_fill_red_black_order_even_dims(shape):
    """With even dimensions every neighbour of a cell has the other colour,
    so one iteration equals: all even cells updated at once from the old
    values, then all odd cells at once from the new ones."""
    ny, nx = shape
    r = np.random.default_rng(5)
    a = r.uniform(-1, 1, size=(3, ny, nx))
    where = r.uniform(size=a.shape) > 0.3
    cx, cy = 1 / (2 / (nx - 1)) ** 2, 1 / (2 / (ny - 1)) ** 2
    dc = 1 / (2 * (cx + cy))
    exp = a.copy()
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    for par in (0, 1):
        xt = cx * (np.roll(exp, 1, axis=2) + np.roll(exp, -1, axis=2))
        yt = cy * (np.roll(exp, 1, axis=1) + np.roll(exp, -1, axis=1))
        upd = (xt + yt) * dc
        m = where & (((ii + jj) & 1) == par)[None]
        exp = np.where(m, upd, exp)
    got = oracle.fill_gauss_seidel_2d(a, where, 1)
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("shape", [(3, 3), (3, 4), (4, 5), (5, 5), (2, 3), (3, 2)])
def test_fill_c_matches_python_loops(shape):
    """The C sweep against the literal pure-Python loops, odd sizes
    included (where the periodic wrap couples cells of one colour)."""
    ny, nx = shape
    r = np.random.default_rng(9)
    a = r.uniform(-1, 1, size=(2, ny, nx))
    where = r.uniform(size=a.shape) > 0.4
    assert np.array_equal(oracle.fill_gauss_seidel_2d(a, where, 3),
                          oracle.fill_gauss_seidel_2d_py(a, where, 3))


SZ_X, SZ_Y = 10, 11


@pytest.mark.parametrize("x", [np.arange(SZ_X), np.linspace(0, SZ_X - 1, num=21)])
@pytest.mark.parametrize("axis", [None, 0, -1])
def test_interp_1d_map_coordinates(x, axis):
    """test_interp_ndarray.py:10-42"""
    a = rng.random(SZ_X)
    result = oracle.ndarray_linear_interpolation(a, (x,), axis=axis)
    expected = scipy.ndimage.map_coordinates(a, x[np.newaxis], mode="nearest", order=1)
    assert np.allclose(result, expected)


@pytest.mark.parametrize("x", [np.arange(SZ_X)[:, None], np.linspace(0, SZ_X - 1, 100)[:, None]])
@pytest.mark.parametrize("y", [np.arange(SZ_Y)[None, :], np.linspace(0, SZ_Y - 1, 5)[None, :]])
@pytest.mark.parametrize("axis", [None, (0, 1), (0, ~0)])
def test_interp_2d_map_coordinates(x, y, axis):
    """test_interp_ndarray.py:45-97"""
    a = rng.random((SZ_X, SZ_Y))
    x, y = np.broadcast_arrays(x, y)
    result = oracle.ndarray_linear_interpolation(a, (x, y), axis=axis)
    expected = scipy.ndimage.map_coordinates(a, np.stack([x, y]), order=1)
    assert np.allclose(result, expected)


def test_interp_linear_functions_reproduced_outside_the_grid():
    """The cell index is clamped, not the value: linear functions are
    extrapolated exactly from the edge cells, bilinear ones likewise."""
    xs = np.linspace(-3.5, 12.5, 41)
    a1 = 3 * np.arange(SZ_X) + 1.0
    np.testing.assert_allclose(oracle.interp_linear_1d(a1[None], xs[None])[0], 3 * xs + 1,
                               rtol=0, atol=1e-12)
    gx, gy = np.meshgrid(np.arange(SZ_X, dtype=float), np.arange(SZ_Y, dtype=float),
                         indexing="ij")
    f = lambda x, y: 0.5 + 2 * x - 3 * y + 0.25 * x * y
    px = rng.uniform(-2, SZ_X + 2, 200)
    py = rng.uniform(-2, SZ_Y + 2, 200)
    got = oracle.interp_linear_2d(f(gx, gy)[None], px[None], py[None])[0]
    np.testing.assert_allclose(got, f(px, py), rtol=0, atol=1e-11)


def test_interp_orthogonal_axes_against_np_interp():
    """Rows of `a` along the orthogonal axis pair with rows of the indices:
    each result row is numpy.interp of its own row (points inside)."""
    a = rng.random((3, 4, SZ_X))
    x = rng.uniform(0, SZ_X - 1, size=(3, 4, 7))
    got = oracle.ndarray_linear_interpolation(a, (x,), axis=-1, axis_indices=-1)
    for k in np.ndindex(3, 4):
        np.testing.assert_allclose(got[k], np.interp(x[k], np.arange(SZ_X), a[k]),
                                   rtol=0, atol=1e-14)
    # the interpolated axis in the middle, indices broadcast over a's rows
    a2 = np.moveaxis(a, -1, 1)
    x2 = x[0, 0][None, :, None]
    got2 = oracle.ndarray_linear_interpolation(a2, (x2,), axis=1, axis_indices=1)
    assert got2.shape == (3, 7, 4)
    for k0 in range(3):
        for k2 in range(4):
            np.testing.assert_allclose(got2[k0, :, k2], np.interp(x2[0, :, 0], np.arange(SZ_X),
                                                                  a2[k0, :, k2]), atol=1e-14)
