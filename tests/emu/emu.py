"""
ctypes front end of the CPU emulation of the GPU build (``librg_emu.so``).

Test infrastructure: it runs the kernels' per-lane bodies on the host so that
the chunk decomposition, the fix-up pass and the bucket reduction can be
compared with the oracle bit for bit without a GPU.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PATH = os.path.join(_HERE, "librg_emu.so")
_lib = None

_pd = ctypes.POINTER(ctypes.c_double)
_pi = ctypes.POINTER(ctypes.c_int64)
_i64 = ctypes.c_int64


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_PATH):
            build()
        L = ctypes.CDLL(_PATH)
        L.rge_build.restype = _i64
        L.rge_build.argtypes = [
            _pd, _pd, _i64, _i64, _pd, _pd, _i64, _i64, _pd, _i64, _i64, ctypes.c_int,
            ctypes.POINTER(_pi), ctypes.POINTER(_pi), ctypes.POINTER(_pd), _pi,
            ctypes.POINTER(_pi), ctypes.POINTER(_pi), ctypes.POINTER(_pd), _pi,
        ]
        L.rge_free.restype = None
        L.rge_free.argtypes = [ctypes.c_void_p]
        L.rge_reduceat_segment.restype = ctypes.c_double
        L.rge_reduceat_segment.argtypes = [_pd, _i64]
        L.rge_cell_area.restype = None
        L.rge_reduceat_segment_inplace.restype = ctypes.c_double
        L.rge_reduceat_segment_inplace.argtypes = [_pd, _i64, _i64]
        L.rge_cell_area.argtypes = [_pd, _pd, _i64, _i64, _pd]
        L.rge_index_secant.restype = _i64
        L.rge_index_secant.argtypes = [_pd, _pd, _i64, _i64, ctypes.c_double, ctypes.c_double]
        L.rge_sort_network.restype = None
        L.rge_sort_network.argtypes = [ctypes.c_void_p, ctypes.c_void_p, _i64, ctypes.c_int]
        _lib = L
    return _lib


def sort_network(keys, n):
    """The kernels' register sorting network over each row of `keys`
    (rows of n = 8 or 16 uint64 keys): the sorted keys and, as values, the
    positions they came from."""
    k = np.ascontiguousarray(keys, dtype=np.uint64).reshape(-1, n).copy()
    v = np.tile(np.arange(n, dtype=np.float64), (k.shape[0], 1))
    lib().rge_sort_network(k.ctypes.data, v.ctypes.data, k.shape[0], int(n))
    return k, v.astype(np.int64)


def _d(a):
    return a.ctypes.data_as(_pd)


def _take(p, n, dtype):
    if n == 0:
        return np.zeros(0, dtype)
    return np.ctypeslib.as_array(p, shape=(n,)).copy()


def build_2d(grid_input, grid_output, weights_input=None, K: int = 16, guess: int = 0,
             stage_per_edge: int = 8):
    """
    Emulated GPU build.  Returns ``(flat, coalesced, stats)``: the
    uncoalesced triples in emission order, the coalesced triples sorted by
    (input, output), and ``dict(chunks, fixups, anomalies, long, failed)``.
    """
    xi = np.ascontiguousarray(grid_input[0], dtype=np.float64)
    yi = np.ascontiguousarray(grid_input[1], dtype=np.float64)
    xo = np.ascontiguousarray(grid_output[0], dtype=np.float64)
    yo = np.ascontiguousarray(grid_output[1], dtype=np.float64)
    wi = None if weights_input is None else np.ascontiguousarray(weights_input, np.float64)
    fi, fo, fw, ci, co, cw = _pi(), _pi(), _pd(), _pi(), _pi(), _pd()
    nf = _i64()
    stats = np.zeros(7, np.int64)
    L = lib()
    nnz = L.rge_build(
        _d(xi), _d(yi), xi.shape[0], xi.shape[1], _d(xo), _d(yo), xo.shape[0], xo.shape[1],
        None if wi is None else _d(wi), K, stage_per_edge, guess,
        ctypes.byref(fi), ctypes.byref(fo), ctypes.byref(fw), ctypes.byref(nf),
        ctypes.byref(ci), ctypes.byref(co), ctypes.byref(cw),
        stats.ctypes.data_as(_pi),
    )
    n = nf.value
    flat = (_take(fi, n, np.int64), _take(fo, n, np.int64), _take(fw, n, np.float64))
    coal = (_take(ci, nnz, np.int64), _take(co, nnz, np.int64), _take(cw, nnz, np.float64))
    for p in (fi, fo, fw, ci, co, cw):
        L.rge_free(ctypes.cast(p, ctypes.c_void_p))
    keys = ("chunks", "fixups", "anomalies", "long", "failed", "restaged", "map_errors")
    return flat, coal, dict(zip(keys, stats.tolist()))


def reduceat_segment(a) -> float:
    a = np.ascontiguousarray(a, dtype=np.float64)
    return float(lib().rge_reduceat_segment(_d(a), a.size))


def bucket_reduce(keys, vals, stride: int = 1):
    """bucket_reduce_sorted on a strided copy of one sorted bucket: returns
    (count, long_run, keys, vals) with the bucket as the function left it."""
    n = len(keys)
    kb = np.full(max(n * stride, 1), np.uint64(0xDEADBEEFDEADBEEF), dtype=np.uint64)
    vb = np.full(max(n * stride, 1), np.nan)
    kb[: n * stride: stride] = keys
    vb[: n * stride: stride] = vals
    lr = ctypes.c_int32(0)
    L = lib()
    L.rge_bucket_reduce.restype = _i64
    L.rge_bucket_reduce.argtypes = [ctypes.c_void_p, _pd, _i64, _i64, ctypes.POINTER(ctypes.c_int32)]
    m = L.rge_bucket_reduce(kb.ctypes.data, _d(vb), n, stride, ctypes.byref(lr))
    return int(m), bool(lr.value), kb[: n * stride: stride].copy(), vb[: n * stride: stride].copy()


def reduceat_segment_inplace(a, stride: int = 1) -> float:
    """The stackless in-place variant (on a strided copy of ``a``)."""
    a = np.asarray(a, dtype=np.float64)
    buf = np.full(a.size * stride, np.nan)
    buf[::stride] = a
    return float(lib().rge_reduceat_segment_inplace(_d(buf), a.size, stride))


def cell_area(x, y) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    out = np.empty((x.shape[0] - 1, x.shape[1] - 1))
    lib().rge_cell_area(_d(x), _d(y), x.shape[0], x.shape[1], _d(out))
    return out
