"""
1-D weights on the GPU (conservative sweep, multilinear) against the oracle,
bit for bit, through the C ABI and through the drop-in; the reference's own
known answers through the drop-in; the block-diagonal apply of all elements
in one launch equals the oracle's per-element apply.
"""

import ctypes

import numpy as np
import pytest

import regridding_amd
from oracle import oracle
from regridding_amd._lib import DeviceArray, check, lib
from tests import oracle_pipeline as op
from tests.test_oracle_line import C1D, WML

pytestmark = pytest.mark.gpu


def _grids(rng, D, n, m, desc_in=False, desc_out=False, lo=-1.3, hi=1.3):
    xi = np.sort(rng.uniform(-1, 1, (D, n)), axis=1)
    xo = np.sort(rng.uniform(lo, hi, (D, m)), axis=1)
    if desc_in:
        xi = xi[:, ::-1]
    if desc_out:
        xo = xo[:, ::-1]
    return np.ascontiguousarray(xi), np.ascontiguousarray(xo)


@pytest.mark.parametrize("desc_in", [False, True])
@pytest.mark.parametrize("desc_out", [False, True])
@pytest.mark.parametrize("with_wi", [False, True])
def test_conservative_1d_c_abi_matches_oracle(desc_in, desc_out, with_wi):
    rng = np.random.default_rng(11 + 2 * desc_in + desc_out)
    D, n, m = 7, 40, 23
    xi, xo = _grids(rng, D, n, m, desc_in, desc_out)
    xo[3] = np.linspace(2, 3, m)[:: -1 if desc_out else 1]  # disjoint element
    xo[4, :] = xi[4, np.linspace(0, n - 1, m).astype(int)]  # shared vertices (ties)
    wi = rng.uniform(0.5, 2, (D, n - 1)) if with_wi else None
    L = lib()
    dxi, dxo = DeviceArray.from_host(xi), DeviceArray.from_host(xo)
    dwi = DeviceArray.from_host(wi) if wi is not None else None
    counts = DeviceArray.empty((D * (n - 1),), np.int32)
    offs = DeviceArray.empty((D * (n - 1) + 1,), np.int64)
    total = ctypes.c_int64()
    check(L.rg_conservative_1d_count(dxi.ptr, n, dxo.ptr, m, D, counts.ptr, offs.ptr,
                                     ctypes.byref(total), 0))
    T = total.value
    ii, io, w = (DeviceArray.empty((max(T, 1),), t) for t in (np.int64, np.int64, np.float64))
    check(L.rg_conservative_1d_emit(dxi.ptr, n, dxo.ptr, m, D, dwi.ptr if dwi else None,
                                    counts.ptr, offs.ptr, ii.ptr, io.ptr, w.ptr, 0))
    gi, go, gw = ii.to_host()[:T], io.to_host()[:T], w.to_host()[:T]
    off = offs.to_host()[:: n - 1]
    for d in range(D):
        ref = oracle.coalesce(*oracle.conservative_1d(xi[d], xo[d], None if wi is None else wi[d]))
        s, e = off[d], off[d + 1]
        assert np.array_equal(gi[s:e], ref[0]), d
        assert np.array_equal(go[s:e], ref[1]), d
        assert np.array_equal(gw[s:e], ref[2]), d


@pytest.mark.parametrize("bounds", ["extrapolate", "nan"])
def test_multilinear_matches_oracle(bounds):
    rng = np.random.default_rng(21)
    xi = np.sort(rng.uniform(-1, 1, (5, 33)), axis=1)
    xo = rng.uniform(-1.4, 1.4, (5, 50))
    xo[:, :3] = xi[:, :3]
    xo[:, 3] = xi[:, -1]
    wi = rng.uniform(0.5, 2, (5, 33))
    got = regridding_amd.weights((xi,), (xo,), axis_input=1, axis_output=1, weights_input=wi,
                                 method="multilinear", bounds=bounds)
    ref = op.weights_multilinear((xi,), (xo,), 1, 1, weights_input=wi, bounds=bounds)
    assert got[1] == ref[1] and got[2] == ref[2]
    for a, b in zip(got[0].reshape(-1), ref[0].reshape(-1)):
        for p, q in zip(a, b):
            assert np.array_equal(p, q, equal_nan=True)


def test_multilinear_bounds_raise():
    with pytest.raises(ValueError, match="fall outside the input grid"):
        regridding_amd.weights(np.linspace(0, 1, 5), np.linspace(-1, 2, 7),
                               method="multilinear", bounds="raise")
    with pytest.raises(ValueError, match="Unrecognized bounds"):
        regridding_amd.weights(np.linspace(0, 1, 5), np.linspace(0, 1, 7),
                               method="multilinear", bounds="clip")
    # the reference searches first (searchsorted), which rejects 2-D grids
    # before the multilinear builder's own check
    x = np.meshgrid(np.linspace(0, 1, 4), np.linspace(0, 1, 4), indexing="ij")
    with pytest.raises(ValueError, match="2-dimensional searchsorted search not supported"):
        regridding_amd.weights(x, x, method="multilinear")


@pytest.mark.parametrize("case", range(len(C1D)))
def test_dropin_conservative_1d_known_answers(case):
    xi, xo, v, expected, ai, ao, wi = C1D[case]
    w = regridding_amd.weights(xi, xo, ai, ao, weights_input=wi, method="conservative")
    got = regridding_amd.regrid_from_weights(*w, values_input=v, axis_input=ai, axis_output=ao)
    assert np.allclose(got, expected)
    ref_w = op.weights_conservative_1d(xi, xo, ai, ao, weights_input=wi)
    for a, b in zip(w[0].reshape(-1), ref_w[0].reshape(-1)):
        for p, q in zip(a, b):
            assert np.array_equal(p, q)
    ref = op.regrid_from_weights(*ref_w, values_input=v, axis_input=ai, axis_output=ao)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("case", range(len(WML)))
def test_dropin_multilinear_known_answers(case):
    ci, co, v, ai, ao, wi, expected = WML[case]
    w = regridding_amd.weights(ci, co, ai, ao, weights_input=wi, method="multilinear")
    got = regridding_amd.regrid_from_weights(*w, values_input=v, axis_input=ai, axis_output=ao)
    assert np.allclose(got, expected)
    ref_w = op.weights_multilinear(ci, co, ai, ao, weights_input=wi)
    ref = op.regrid_from_weights(*ref_w, values_input=v, axis_input=ai, axis_output=ao)
    assert np.array_equal(got, ref)


def test_regrid_default_method_is_multilinear():
    x = np.linspace(-1, 1, 21)
    xo = np.linspace(-0.7, 0.9, 13)
    got = regridding_amd.regrid(x, xo, 2 * x + 1)
    assert np.allclose(got, 2 * xo + 1, atol=1e-13)


def test_block_diagonal_apply_and_transpose():
    rng = np.random.default_rng(31)
    xi, xo = _grids(rng, 12, 30, 18, desc_in=True)
    w = regridding_amd.weights((xi,), (xo,), axis_input=1, axis_output=1, method="conservative")
    v = rng.normal(size=(12, 29))
    got = regridding_amd.regrid_from_weights(*w, values_input=v, axis_input=1, axis_output=1)
    ref = op.regrid_from_weights(*w, values_input=v, axis_input=1, axis_output=1)
    assert np.array_equal(got, ref)
    wt = regridding_amd.transpose_weights(w)
    y = rng.normal(size=(12, 17))
    got_t = regridding_amd.regrid_from_weights(*wt, values_input=y, axis_input=1, axis_output=1)
    ref_t = op.regrid_from_weights(*wt, values_input=y, axis_input=1, axis_output=1)
    assert np.array_equal(got_t, ref_t)
