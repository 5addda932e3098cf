"""
``fill`` (red-black Gauss-Seidel) and ``ndarray_linear_interpolation`` on
the GPU against the oracle, bit for bit, through the C ABI and through the
drop-in; the reference's own tests (``_fill/_fill_test.py``,
``_tests/test_interp_ndarray.py``) through the drop-in; error behaviour.

The reference evaluates ``_bilinear_interpolation`` and the Gauss-Seidel
cell update with numba ``fastmath=True``; its low bits are not pinned (the
GPU and the oracle both compute the written order with IEEE rounding).
"""

import numpy as np
import pytest
import scipy.ndimage

import regridding_amd
from oracle import oracle
from regridding_amd._lib import DeviceArray, check, lib
from tests.test_oracle_fill_interp import NT, NX, NY, _fill_cases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(1, 4, 6), (3, 5, 7), (2, 7, 4), (4, 12, 11), (2, 11, 12),
                                   (3, 9, 9), (2, 2, 3), (5, 2, 2), (1, 65, 130), (2, 97, 61)])
@pytest.mark.parametrize("iters", [1, 7])
def test_fill_c_abi_matches_oracle(shape, iters):
    T, ny, nx = shape
    r = np.random.default_rng(ny * 1000 + nx + iters)
    a = r.uniform(-1, 1, size=shape)
    where = r.uniform(size=shape) > 0.5
    da = DeviceArray.from_host(a)
    dw = DeviceArray.from_host(where.astype(np.uint8))
    check(lib().rg_fill_gauss_seidel_2d(da.ptr, dw.ptr, T, ny, nx, iters, 0))
    assert np.array_equal(da.to_host(), oracle.fill_gauss_seidel_2d(a, where, iters))


@pytest.mark.parametrize("case", range(3))
@pytest.mark.parametrize("guess", [None, 0.5])
def test_fill_reference_cases_match_oracle(case, guess):
    """_fill_test.py:11-57 through the drop-in, compared bit for bit."""
    a, where, axis = _fill_cases()[case]
    got = regridding_amd.fill(a, where, axis, guess=guess, num_iterations=11)
    exp = oracle.fill(a, where, axis, guess=guess, num_iterations=11)
    assert np.array_equal(got, exp)
    if where is None:
        where = np.isnan(a)
    assert np.all(np.isfinite(got))
    assert np.allclose(got[~where], a[~where])
    assert np.all(got[where] != 0)


@pytest.mark.parametrize("guess", [None, 2.0, np.arange(NT).reshape(NT, 1, 1)])
def test_fill_guess_zero_iterations(guess):
    r = np.random.default_rng(3)
    a = r.uniform(0, 1, size=(NT, NX, NY))
    where = r.uniform(0, 1, size=a.shape) > 0.9
    result = regridding_amd.fill(a, where, axis=(~1, ~0), guess=guess, num_iterations=0)
    if guess is None:
        guess = np.nanmedian(np.where(where, np.nan, a), axis=(~1, ~0), keepdims=True)
    assert np.allclose(result[~where], a[~where])
    assert np.allclose(result[where], np.broadcast_to(guess, a.shape)[where])


def test_fill_missing_cluster_and_all_missing():
    x = np.linspace(-1, 1, num=32)
    a = x[:, None] + 2 * x[None, :]
    am = a.copy()
    am[10:16, 10:16] = np.nan
    got = regridding_amd.fill(am, num_iterations=1000)
    assert np.array_equal(got, oracle.fill(am, num_iterations=1000))
    assert np.allclose(got, a, atol=1e-4)

    r = np.random.default_rng(4)
    b = r.uniform(0, 1, size=(NT, NX, NY))
    where = r.uniform(0, 1, size=b.shape) > 0.9
    where[0] = True
    got = regridding_amd.fill(b, where, axis=(~1, ~0), num_iterations=11)
    assert np.all(got[0] == 0)
    assert np.array_equal(got, oracle.fill(b, where, axis=(~1, ~0), num_iterations=11))


def test_fill_axes_in_the_middle_and_reversed():
    r = np.random.default_rng(8)
    a = r.uniform(0, 1, size=(5, 9, 3, 8))
    where = r.uniform(size=(9, 1, 8)) > 0.7      # broadcast against a
    for axis in [(1, 3), (3, 1), (-1, -3), (0, 2)]:
        got = regridding_amd.fill(a, where, axis, num_iterations=5)
        assert np.array_equal(got, oracle.fill(a, where, axis, num_iterations=5)), axis


def test_fill_errors():
    a = np.random.default_rng(0).uniform(size=(4, 5, 6))
    with pytest.raises(ValueError):
        regridding_amd.fill(a, a > 0.5, axis=(0, 1, 2))
    with pytest.raises(ZeroDivisionError):
        regridding_amd.fill(a[:, :1], a[:, :1] > 0.5, axis=(0, 1))
    assert np.array_equal(regridding_amd.fill(a[:, :1], a[:, :1] > 0.5, axis=(0, 1),
                                              num_iterations=0, guess=1.0),
                          np.where(a[:, :1] > 0.5, 1.0, a[:, :1]))


def test_fill_large_slices():
    """Many slices of a larger grid: the parallel colour kernel and the
    late-cell kernel across slice boundaries (odd nx, odd ny)."""
    r = np.random.default_rng(21)
    for shape in [(64, 129, 257), (16, 256, 255)]:
        a = r.uniform(-1, 1, size=shape)
        where = r.uniform(size=shape) > 0.8
        got = regridding_amd.fill(a, where, axis=(1, 2), num_iterations=3)
        assert np.array_equal(got, oracle.fill(a, where, (1, 2), num_iterations=3)), shape


SZ_X, SZ_Y = 10, 11


@pytest.mark.parametrize("x", [np.arange(SZ_X), np.linspace(0, SZ_X - 1, num=21),
                               np.linspace(-3, SZ_X + 2, 40)])
@pytest.mark.parametrize("axis", [None, 0, -1])
def test_interp_1d(x, axis):
    a = np.random.default_rng(1).random(SZ_X)
    got = regridding_amd.ndarray_linear_interpolation(a, (x,), axis=axis)
    assert np.array_equal(got, oracle.ndarray_linear_interpolation(a, (x,), axis=axis))
    if x.min() >= 0 and x.max() <= SZ_X - 1:
        exp = scipy.ndimage.map_coordinates(a, x[np.newaxis], mode="nearest", order=1)
        assert np.allclose(got, exp)


@pytest.mark.parametrize("x", [np.arange(SZ_X)[:, None], np.linspace(0, SZ_X - 1, 100)[:, None],
                               np.linspace(-2, SZ_X + 1, 30)[:, None]])
@pytest.mark.parametrize("y", [np.arange(SZ_Y)[None, :], np.linspace(0, SZ_Y - 1, 5)[None, :]])
@pytest.mark.parametrize("axis", [None, (0, 1), (0, ~0), (~0, 0)])
def test_interp_2d(x, y, axis):
    a = np.random.default_rng(2).random((SZ_X, SZ_Y))
    x, y = np.broadcast_arrays(x, y)
    got = regridding_amd.ndarray_linear_interpolation(a, (x, y), axis=axis)
    assert np.array_equal(got, oracle.ndarray_linear_interpolation(a, (x, y), axis=axis))
    if x.min() >= 0 and x.max() <= SZ_X - 1:
        exp = scipy.ndimage.map_coordinates(a, np.stack([x, y]), order=1)
        assert np.allclose(got, exp)


@pytest.mark.parametrize("axis,axis_indices,a_shape,i_shape", [
    (-1, -1, (3, 4, SZ_X), (3, 4, 7)),
    (1, 1, (3, SZ_X, 4), (1, 7, 1)),
    (0, 0, (SZ_X, 2, 3), (7, 1, 3)),
    (-1, 0, (2, 3, SZ_X), (7, 2, 3)),
    ((0, 2), (0, 1), (SZ_X, 3, SZ_Y), (5, 6, 3)),
    ((2, 0), (1, 0), (SZ_X, 3, SZ_Y), (5, 6, 1)),
    ((1, 2), (1, 2), (4, SZ_X, SZ_Y), (4, 6, 2)),
])
def test_interp_broadcasting(axis, axis_indices, a_shape, i_shape):
    r = np.random.default_rng(len(a_shape) * 10 + len(i_shape))
    a = r.random(a_shape)
    naxes = 1 if isinstance(axis, int) else len(axis)
    idx = tuple(r.uniform(-1, 12, size=i_shape) for _ in range(naxes))
    got = regridding_amd.ndarray_linear_interpolation(a, idx, axis=axis, axis_indices=axis_indices)
    exp = oracle.ndarray_linear_interpolation(a, idx, axis=axis, axis_indices=axis_indices)
    assert got.shape == exp.shape
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("axis,axis_indices,a_shape,i_shape", [
    (0, None, (SZ_X, 2, 3), (7,)),
    ((1, 2), None, (4, SZ_X, SZ_Y), (6, 2)),
])
def test_interp_fewer_orthogonal_index_axes_raise_like_the_reference(axis, axis_indices,
                                                                     a_shape, i_shape):
    """The reference places the broadcast orthogonal axes by position in
    each operand's own orthogonal axes, so indices with fewer orthogonal
    axes than `a` raise ValueError (_interp_ndarray.py:112-124)."""
    r = np.random.default_rng(0)
    a = r.random(a_shape)
    naxes = 1 if isinstance(axis, int) else len(axis)
    idx = tuple(r.uniform(0, 9, size=i_shape) for _ in range(naxes))
    with pytest.raises(ValueError):
        oracle.ndarray_linear_interpolation(a, idx, axis=axis, axis_indices=axis_indices)
    with pytest.raises(ValueError):
        regridding_amd.ndarray_linear_interpolation(a, idx, axis=axis,
                                                    axis_indices=axis_indices)


def test_interp_large_and_errors():
    r = np.random.default_rng(5)
    a = r.random((32, 700, 300))
    x = r.uniform(-5, 705, size=(32, 4096))
    y = r.uniform(-5, 305, size=(32, 4096))
    got = regridding_amd.ndarray_linear_interpolation(a, (x, y), axis=(1, 2), axis_indices=(1,))
    assert np.array_equal(got, oracle.interp_linear_2d(a, x, y))
    with pytest.raises(ValueError):
        regridding_amd.ndarray_linear_interpolation(a, (x,), axis=(1, 2))
    with pytest.raises(NotImplementedError):
        regridding_amd.ndarray_linear_interpolation(a, (x, x, x))
