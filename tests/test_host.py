"""
Host-side checks that need no GPU: the C ABI library loads and exports every
function declared in include/regridding_amd.h, the ctypes signatures cover
them, and the drop-in's argument handling raises the reference's errors
before any device call.
"""

import ctypes
import os
import re

import numpy as np
import pytest

import regridding_amd
from regridding_amd import _lib, _util

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "regridding_amd.h")).read()
    # function declarations: a return type, the name, an opening parenthesis
    decl = r"^(?:const\s+char\s*\*|int64_t|int|void)\s*(rg_[a-z0-9_]+)\s*\("
    return sorted(set(re.findall(decl, text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    L = _lib.load_library()
    names = _declared()
    assert len(names) >= 25
    for name in names:
        assert hasattr(L, name), name
        assert isinstance(getattr(L, name), ctypes._CFuncPtr)


def test_ctypes_signatures_cover_the_header():
    assert sorted(_lib.SIGNATURES) == _declared()


def test_normalize_axis():
    assert _util.normalize_axis(None, 3) == (-3, -2, -1)
    assert _util.normalize_axis(0, 2) == (-2,)
    assert _util.normalize_axis((1, 0), 2) == (-1, -2)
    with pytest.raises(np.exceptions.AxisError):
        _util.normalize_axis(3, 2)


def test_normalize_errors_match_reference():
    x = np.zeros((4, 5))
    with pytest.raises(ValueError, match="must match the number of axes"):
        _util.normalize_coordinates((x, x), (x, x), axis_input=(0, 1), axis_output=0)
    with pytest.raises(ValueError, match="should match the number of axes"):
        _util.normalize_coordinates((x,), (x,), axis_input=(0, 1), axis_output=(0, 1))
    with pytest.raises(ValueError, match="should match the number of elements"):
        _util.normalize_coordinates((x,), (x, x), axis_input=0, axis_output=0)


def test_normalize_broadcasts_orthogonal_axes():
    xi = np.zeros((10, 11, 2))
    xo = np.zeros((7, 8, 1))
    ci, co, ai, ao, si, so, orth = _util.normalize_coordinates(
        (xi, xi), (xo, xo), axis_input=(0, 1), axis_output=(0, 1))
    assert ai == (-2, -3) and ao == (-2, -3)
    assert orth == (2,)
    assert ci[0].shape == (10, 11, 2) and co[0].shape == (7, 8, 2)


def test_weights_rejects_unknown_method_before_device():
    x = np.meshgrid(np.linspace(0, 1, 3), np.linspace(0, 1, 3), indexing="ij")
    with pytest.raises(ValueError, match="unrecognized method"):
        regridding_amd.weights(x, x, method="bogus")


def test_find_indices_2d_reference_methods_raise():
    x = np.meshgrid(np.linspace(0, 1, 3), np.linspace(0, 1, 3), indexing="ij")
    with pytest.raises(ValueError, match="2-dimensional brute-force search not supported"):
        regridding_amd.find_indices(x, x, method="brute")
    with pytest.raises(ValueError, match="2-dimensional searchsorted search not supported"):
        regridding_amd.find_indices(x, x, method="searchsorted")
    with pytest.raises(ValueError, match="not recognized"):
        regridding_amd.find_indices((x[0][:, 0],), (x[0][:, 0],), method="invalid method")


def test_transpose_weights_is_a_view_swap():
    ii, io, w = np.arange(3), np.arange(3) + 1, np.ones(3)
    arr = np.empty((), dtype=object)
    arr[()] = (ii, io, w)
    t, si, so = regridding_amd.transpose_weights((arr, (4,), (5,)))
    assert (si, so) == ((5,), (4,))
    assert t[()][0] is io and t[()][1] is ii and t[()][2] is w
