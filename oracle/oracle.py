"""
ctypes front end of the CPU oracle (``oracle/librg_oracle.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` -- never by ``regridding_amd``.
The C source (``rg_oracle.c``) restates the reference's conservative 2-D path
function by function; this module only marshals numpy arrays.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "librg_oracle.so")

_i64 = ctypes.c_int64
_dbl = ctypes.c_double
_pd = ctypes.POINTER(ctypes.c_double)
_pi = ctypes.POINTER(ctypes.c_int64)

_lib = None


def build() -> None:
    """Compile the oracle with its Makefile (gcc, IEEE exact order)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.rgo_bbox_intersect.restype = ctypes.c_int
        L.rgo_bbox_intersect.argtypes = [_dbl] * 8
        L.rgo_seg_params.restype = None
        L.rgo_seg_params.argtypes = [_pd, _pd]
        L.rgo_seg_hit.restype = ctypes.c_int
        L.rgo_seg_hit.argtypes = [_dbl, _dbl]
        L.rgo_point_in_polygon.restype = ctypes.c_int
        L.rgo_point_in_polygon.argtypes = [_dbl, _dbl, _pd, _pd, _i64]
        L.rgo_grid_volume.restype = None
        L.rgo_grid_volume.argtypes = [_pd, _pd, _i64, _i64, _pd]
        L.rgo_grid_boundary.restype = _i64
        L.rgo_grid_boundary.argtypes = [_pd, _pd, _i64, _i64, _pd, _pd]
        for name in ("rgo_index_of_point_brute", "rgo_index_of_point_secant"):
            f = getattr(L, name)
            f.restype = None
            f.argtypes = [_dbl, _dbl, _pd, _pd, _i64, _i64, _pi]
        L.rgo_sweep_2d.restype = _i64
        L.rgo_sweep_2d.argtypes = [
            _pd, _pd, _i64, _i64, _pd, _pd, _i64, _i64, _pd,
            ctypes.POINTER(_pi), ctypes.POINTER(_pi), ctypes.POINTER(_pd),
        ]
        L.rgo_free.restype = None
        L.rgo_free.argtypes = [ctypes.c_void_p]
        L.rgo_reduce_segment.restype = _dbl
        L.rgo_reduce_segment.argtypes = [_pd, _i64]
        L.rgo_coalesce.restype = _i64
        L.rgo_coalesce.argtypes = [_i64, _pi, _pi, _pd, _pi, _pi, _pd]
        L.rgo_apply.restype = None
        L.rgo_apply.argtypes = [_i64, _pi, _pi, _pd, _i64, _pd, _i64, _pd, _i64]
        for name in ("rgo_find_indices_brute_1d", "rgo_find_indices_searchsorted_1d"):
            f = getattr(L, name)
            f.restype = None
            f.argtypes = [_pd, _i64, _pd, _i64, _i64, _i64, _pi]
        L.rgo_fill_gauss_seidel_2d.restype = None
        L.rgo_fill_gauss_seidel_2d.argtypes = [_pd, ctypes.c_void_p, _i64, _i64, _i64, _i64]
        L.rgo_interp_linear_1d.restype = None
        L.rgo_interp_linear_1d.argtypes = [_pd, _i64, _pd, _i64, _i64, _pd]
        L.rgo_interp_linear_2d.restype = None
        L.rgo_interp_linear_2d.argtypes = [_pd, _i64, _i64, _pd, _pd, _i64, _i64, _pd]
        L.rgo_set_threads.restype = None
        L.rgo_set_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def _d(a: np.ndarray):
    return a.ctypes.data_as(_pd)


def _i(a: np.ndarray):
    return a.ctypes.data_as(_pi)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def set_threads(n: int) -> None:
    lib().rgo_set_threads(int(n))


def bbox_intersect(*v: float) -> bool:
    return bool(lib().rgo_bbox_intersect(*[float(q) for q in v]))


def seg_params(p1, p2, q1, q2) -> tuple[float, float]:
    pq = _f64([*p1, *p2, *q1, *q2])
    tu = np.zeros(2)
    lib().rgo_seg_params(_d(pq), _d(tu))
    return float(tu[0]), float(tu[1])


def seg_hit(t: float, u: float) -> bool:
    return bool(lib().rgo_seg_hit(float(t), float(u)))


def point_in_polygon(x: float, y: float, vx, vy) -> bool:
    vx = _f64(vx)
    vy = _f64(vy)
    return bool(lib().rgo_point_in_polygon(float(x), float(y), _d(vx), _d(vy), vx.size))


def grid_volume(x, y) -> np.ndarray:
    x = _f64(x)
    y = _f64(y)
    nx, ny = x.shape
    out = np.empty((nx - 1, ny - 1))
    lib().rgo_grid_volume(_d(x), _d(y), nx, ny, _d(out))
    return out


def grid_boundary(x, y) -> tuple[np.ndarray, np.ndarray]:
    x = _f64(x)
    y = _f64(y)
    nx, ny = x.shape
    n = 2 * (nx - 1) + 2 * (ny - 1)
    bx = np.empty(n)
    by = np.empty(n)
    lib().rgo_grid_boundary(_d(x), _d(y), nx, ny, _d(bx), _d(by))
    return bx, by


def _index_of_point(name: str, point, grid) -> tuple[int, int]:
    x = _f64(grid[0])
    y = _f64(grid[1])
    nx, ny = x.shape
    ij = np.zeros(2, dtype=np.int64)
    getattr(lib(), name)(float(point[0]), float(point[1]), _d(x), _d(y), nx, ny, _i(ij))
    return int(ij[0]), int(ij[1])


def index_of_point_brute(point, grid) -> tuple[int, int]:
    return _index_of_point("rgo_index_of_point_brute", point, grid)


def index_of_point_secant(point, grid) -> tuple[int, int]:
    return _index_of_point("rgo_index_of_point_secant", point, grid)


def sweep_2d(grid_input, grid_output, weights_input=None):
    """Uncoalesced ``(in, out, w)`` triples in the reference's emission order."""
    xi = _f64(grid_input[0])
    yi = _f64(grid_input[1])
    xo = _f64(grid_output[0])
    yo = _f64(grid_output[1])
    nxi, nyi = xi.shape
    nxo, nyo = xo.shape
    wi = None if weights_input is None else _f64(weights_input)
    pin = _pi()
    pout = _pi()
    pw = _pd()
    L = lib()
    n = L.rgo_sweep_2d(
        _d(xi), _d(yi), nxi, nyi, _d(xo), _d(yo), nxo, nyo,
        None if wi is None else _d(wi),
        ctypes.byref(pin), ctypes.byref(pout), ctypes.byref(pw),
    )
    try:
        if n == 0:
            return (np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0))
        a = np.ctypeslib.as_array(pin, shape=(n,)).copy()
        b = np.ctypeslib.as_array(pout, shape=(n,)).copy()
        c = np.ctypeslib.as_array(pw, shape=(n,)).copy()
    finally:
        L.rgo_free(ctypes.cast(pin, ctypes.c_void_p))
        L.rgo_free(ctypes.cast(pout, ctypes.c_void_p))
        L.rgo_free(ctypes.cast(pw, ctypes.c_void_p))
    return a, b, c


def reduce_segment(a) -> float:
    a = _f64(a)
    return float(lib().rgo_reduce_segment(_d(a), a.size))


def coalesce(indices_input, indices_output, values):
    i = np.ascontiguousarray(indices_input, dtype=np.int64)
    o = np.ascontiguousarray(indices_output, dtype=np.int64)
    w = _f64(values)
    n = i.size
    ic = np.empty(n, np.int64)
    oc = np.empty(n, np.int64)
    wc = np.empty(n)
    m = lib().rgo_coalesce(n, _i(i), _i(o), _d(w), _i(ic), _i(oc), _d(wc))
    return ic[:m].copy(), oc[:m].copy(), wc[:m].copy()


def weights_2d(grid_input, grid_output, weights_input=None):
    """Sweep + coalesce for one 2-D element: the canonical weight triple."""
    return coalesce(*sweep_2d(grid_input, grid_output, weights_input))


def apply(indices_input, indices_output, values, values_input, n_out: int):
    """
    ``values_input`` is ``(D, n_in)``; returns ``(D, n_out)`` float64 built by
    the reference's sequential scatter in storage order.
    """
    i = np.ascontiguousarray(indices_input, dtype=np.int64)
    o = np.ascontiguousarray(indices_output, dtype=np.int64)
    w = _f64(values)
    x = _f64(values_input)
    D, n_in = x.shape
    y = np.zeros((D, n_out))
    lib().rgo_apply(w.size, _i(i), _i(o), _d(w), D, _d(x), n_in, _d(y), n_out)
    return y


def find_indices_1d(x_input, x_output, fill_value: int, method: str):
    """``x_input`` is ``(D, m)``, ``x_output`` ``(D, n)``; returns ``(D, n)``."""
    a = _f64(x_input)
    v = _f64(x_output)
    D, m = a.shape
    _, n = v.shape
    res = np.empty((D, n), np.int64)
    name = {
        "brute": "rgo_find_indices_brute_1d",
        "searchsorted": "rgo_find_indices_searchsorted_1d",
    }[method]
    getattr(lib(), name)(_d(a), m, _d(v), n, D, int(fill_value), _i(res))
    return res


# ---------------------------------------------------------------------------
# 1-D weights (pure Python restatements: the sweeps are O(n + m) per element)
# ---------------------------------------------------------------------------

_MAXSIZE = np.iinfo(np.int64).max  # sys.maxsize


def _index_of_point_1d(point: float, grid: np.ndarray) -> int:
    """_weights_conservative_1d/_grids.py:40-73: bisection, right index."""
    lo, hi = 0, grid.shape[0]
    while hi - lo > 1:
        mid = (lo + hi) // 2
        if grid[mid] > point:
            hi = mid
        else:
            lo = mid
    return hi


def conservative_1d(x_input, x_output, weights_input=None):
    """
    One element of the 1-D conservative sweep, statement by statement after
    ``_weights_conservative_1d.py:57-318``: the uncoalesced triples in the
    reference's emission order, with its negative (wraparound) indices for
    descending grids.
    """
    x_sweep = np.asarray(x_input, dtype=np.float64)
    x_static = np.asarray(x_output, dtype=np.float64)
    num_sweep = x_sweep.shape[0]
    num_static = x_static.shape[0]
    reversed_sweep = not (x_sweep[0] < x_sweep[-1])
    if reversed_sweep:
        x_sweep = x_sweep[::-1]
    reversed_static = not (x_static[0] < x_static[-1])
    if reversed_static:
        x_static = x_static[::-1]
    x_static_left, x_static_right = x_static[0], x_static[-1]
    length_input = x_sweep[1:] - x_sweep[:-1]
    out = []
    index_sweep = 0
    point_1 = x_sweep[0]
    if x_static_left == point_1:
        outside, index_static = False, 0
    elif x_static_left < point_1 < x_static_right:
        outside, index_static = False, _index_of_point_1d(point_1, x_static) - 1
    else:
        outside, index_static = True, _MAXSIZE
    while index_sweep < num_sweep - 1:
        point_2 = x_sweep[index_sweep + 1]
        if outside:  # _step_outside_static, :193-240
            point_static = x_static[0]
            if point_1 < point_static < point_2:
                index_static = 0
                point_2 = point_static
            elif point_static == point_2:
                index_sweep += 1
                index_static = 0
            else:
                index_sweep += 1
            if index_static < _MAXSIZE:
                outside = False
        else:  # _step_inside_static, :243-318
            index_input = ~index_sweep if reversed_sweep else index_sweep
            index_output = ~index_static if reversed_static else index_static
            point_static = x_static[index_static + 1]
            if point_1 < point_static < point_2:
                index_static += 1
                point_2 = point_static
            elif point_static == point_2:
                index_static += 1
                index_sweep += 1
            else:
                index_sweep += 1
            length = point_2 - point_1
            ratio = length / length_input[index_input]
            if weights_input is not None:
                ratio = ratio * weights_input[index_input]
            out.append((index_input, index_output, ratio))
            if not (0 <= index_static < num_static - 1):
                break
        point_1 = point_2
    if not out:
        return np.empty(0, np.int64), np.empty(0, np.int64), np.empty(0)
    ii, io, w = zip(*out)
    return np.array(ii, np.int64), np.array(io, np.int64), np.array(w, np.float64)


def multilinear_1d(indices, x_input, x_output, weights_input=None, bounds_nan=False):
    """
    One element of ``_weights_multilinear.py:95-118`` (index clamping) and
    ``:148-206`` (two weights per output vertex, storage order 2i, 2i+1),
    from the searchsorted indices of the output vertices.
    """
    x_input = np.asarray(x_input, dtype=np.float64)
    x_output = np.asarray(x_output, dtype=np.float64)
    index = np.asarray(indices, dtype=np.int64)
    index_max = x_input.shape[0] - 2
    outside = (index < 0) | (index > index_max)
    below = x_output < x_input[:1]
    i0 = np.where(below, 0, np.clip(index, 0, index_max))
    n = x_output.shape[0]
    ii = np.empty(2 * n, np.int64)
    io = np.empty(2 * n, np.int64)
    w = np.empty(2 * n)
    for i in range(n):
        a = int(i0[i])
        x0, x1, x = x_input[a], x_input[a + 1], x_output[i]
        w1 = (x - x0) / (x1 - x0)
        w0 = 1 - w1
        if weights_input is not None:
            w0 = w0 * weights_input[a]
            w1 = w1 * weights_input[a + 1]
        if bounds_nan and outside[i]:
            w0 = w1 = np.nan
        ii[2 * i], io[2 * i], w[2 * i] = a, i, w0
        ii[2 * i + 1], io[2 * i + 1], w[2 * i + 1] = a + 1, i, w1
    return ii, io, w, int(outside.sum())


def fill_gauss_seidel_2d(a, where, num_iterations: int) -> np.ndarray:
    """``_fill_gauss_seidel_2d`` (``_fill/_gauss_seidel.py:83-139``) on
    (T, ny, nx) arrays: the sequential red-black sweep, in place on a copy."""
    a = np.array(a, dtype=np.float64, order="C", copy=True)
    w = np.ascontiguousarray(where, dtype=np.uint8)
    T, ny, nx = a.shape
    lib().rgo_fill_gauss_seidel_2d(_d(a), w.ctypes.data, T, ny, nx, int(num_iterations))
    return a


def fill_gauss_seidel_2d_py(a, where, num_iterations: int) -> np.ndarray:
    """Pure-Python loops of the same sweep, for tiny inputs."""
    a = np.array(a, dtype=np.float64, copy=True)
    T, ny, nx = a.shape
    dx, dy = 2 / (nx - 1), 2 / (ny - 1)
    dxxinv, dyyinv = 1 / (dx * dx), 1 / (dy * dy)
    dcent = 1 / (2 * (dxxinv + dyyinv))
    for t in range(T):
        for _ in range(num_iterations):
            for is_odd in (0, 1):
                for j in range(ny):
                    for i in range(nx):
                        if (i + j) & 1 == is_odd and where[t, j, i]:
                            xterm = dxxinv * (a[t, j, (i - 1) % nx] + a[t, j, (i + 1) % nx])
                            yterm = dyyinv * (a[t, (j - 1) % ny, i] + a[t, (j + 1) % ny, i])
                            a[t, j, i] = (xterm + yterm) * dcent
    return a


def fill(a, where=None, axis=None, guess=None, num_iterations: int = 100) -> np.ndarray:
    """``fill_gauss_seidel`` (``_fill/_gauss_seidel.py:12-80``) with the
    sweep above."""
    if where is None:
        where = np.isnan(a)
    a = np.array(a, copy=True)
    a, where = np.broadcast_arrays(a, where, subok=True)
    a = np.array(a)
    axis = tuple(range(a.ndim)) if axis is None else np.atleast_1d(axis).tolist()
    axis = tuple(ax % a.ndim for ax in axis)
    if guess is None:
        b = np.where(where, np.nan, a)
        import warnings
        with warnings.catch_warnings():
            warnings.simplefilter("ignore", category=RuntimeWarning)
            med = np.nanmedian(b, axis=axis, keepdims=True)
        guess = np.where(np.isnan(med), 0, med)
    a[where] = np.broadcast_to(guess, a.shape)[where]
    last = (-2, -1)
    am = np.moveaxis(a, axis, last)
    wm = np.moveaxis(where, axis, last)
    shp = am.shape
    r = fill_gauss_seidel_2d(am.reshape(-1, *shp[-2:]), wm.reshape(-1, *shp[-2:]),
                             num_iterations)
    return np.moveaxis(r.reshape(shp), last, axis)


def interp_linear_1d(a, x) -> np.ndarray:
    """``_linear_interpolation`` (``_interp_ndarray.py:208-224``) for every
    row of ``a`` (D, n) at the points ``x`` (D, m)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    D, n = a.shape
    m = x.shape[1]
    out = np.empty((D, m))
    lib().rgo_interp_linear_1d(_d(a), n, _d(x), m, D, _d(out))
    return out


def interp_linear_2d(a, x, y) -> np.ndarray:
    """``_bilinear_interpolation`` (``_interp_ndarray.py:227-269``) for every
    array of ``a`` (D, nx, ny) at the points ``x``, ``y`` (D, m)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    D, nx, ny = a.shape
    m = x.shape[1]
    out = np.empty((D, m))
    lib().rgo_interp_linear_2d(_d(a), nx, ny, _d(x), _d(y), m, D, _d(out))
    return out


def ndarray_linear_interpolation(a, indices, axis=None, axis_indices=None) -> np.ndarray:
    """``ndarray_linear_interpolation`` (``_interp_ndarray.py:11-165``): the
    reference's per-orthogonal-index loop around the two evaluators."""
    a = np.asarray(a)
    shape_a, ndim_a = a.shape, a.ndim
    bi = np.broadcast(*indices)
    shape_indices, ndim_indices = bi.shape, bi.ndim
    if axis is None:
        axis = tuple(range(ndim_a))
    axis = np.lib.array_utils.normalize_axis_tuple(axis, ndim=ndim_a)
    axis = tuple(~(~np.array(axis) % ndim_a))
    if axis_indices is None:
        axis_indices = tuple(range(ndim_indices))
    axis_indices = np.lib.array_utils.normalize_axis_tuple(axis_indices, ndim=ndim_indices)
    axis_indices = tuple(~(~np.array(axis_indices) % ndim_indices))
    if len(indices) != len(axis):
        raise ValueError("The number of indices must match the number of elements in axis")
    oa = tuple(ax for ax in range(-ndim_a, 0) if ax not in axis)
    oi = tuple(ax for ax in range(-ndim_indices, 0) if ax not in axis_indices)
    so = np.broadcast_shapes(tuple(shape_a[ax] for ax in oa),
                             tuple(shape_indices[ax] for ax in oi))
    nba, nbi = len(so) + len(axis), len(so) + len(axis_indices)
    sba = tuple(shape_a[ax] if ax in axis else so[oa.index(ax)] for ax in range(-nba, 0))
    sbi = tuple(shape_indices[ax] if ax in axis_indices else so[oi.index(ax)]
                for ax in range(-nbi, 0))
    a = np.broadcast_to(a, sba)
    indices = tuple(np.broadcast_to(ind, sbi) for ind in indices)
    result = np.empty(sbi)
    for index in np.ndindex(*so):
        ia = tuple(slice(None) if ax in axis else index[oa.index(ax)] for ax in range(-nba, 0))
        ii = tuple(slice(None) if ax in axis_indices else index[oi.index(ax)]
                   for ax in range(-nbi, 0))
        sub = a[ia]
        shp = indices[0][ii].shape
        if len(axis) == 1:
            r = interp_linear_1d(sub[None], indices[0][ii].reshape(1, -1))
        elif len(axis) == 2:
            r = interp_linear_2d(sub[None], indices[0][ii].reshape(1, -1),
                                 indices[1][ii].reshape(1, -1))
        else:
            raise NotImplementedError
        result[ii] = r.reshape(shp)
    return result
