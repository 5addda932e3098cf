"""
Serving loop: continuous batching of regrid requests on one GPU.

``Server`` wraps the native engine (``rg_serve_*`` in ``librg_hip.so``):
requests -- a grid pair (or the key of weights built earlier) and a stack of
value arrays -- are queued in arrival order; worker threads, one stream
each, take the longest prefix of the queue that fits the free device memory,
build all cache misses of that group in one batched build, apply the whole
group in one batched apply per precision and copy the results back.  Built
weights stay on the device, keyed by ``grid_key`` (a hash of the perturbed
coordinates and the input weighting).
"""

from __future__ import annotations

import ctypes

import numpy as np

from ._lib import (RGError, check, lib, rg_serve_config, rg_serve_request, rg_serve_stats,
                   rg_tuning)

__all__ = ["Server", "Ticket", "grid_key", "pinned_empty"]

RG_F64, RG_F32 = 0, 1


def _bind():
    return lib()


def clock_ns() -> int:
    """The engine's monotonic clock (arrival and completion times)."""
    return int(_bind().rg_clock_ns())


class _Pinned:
    def __init__(self, nbytes: int):
        L = _bind()
        self.ptr = ctypes.c_void_p()
        check(L.rg_host_alloc(ctypes.byref(self.ptr), max(int(nbytes), 1)))

    def __del__(self):
        if getattr(self, "ptr", None) and self.ptr.value:
            _bind().rg_host_free(self.ptr)
            self.ptr = None


def pinned_empty(shape, dtype) -> np.ndarray:
    """A numpy array in page-locked host memory (fast asynchronous copies)."""
    dtype = np.dtype(dtype)
    n = int(np.prod(shape)) * dtype.itemsize
    owner = _Pinned(n)
    buf = (ctypes.c_char * max(n, 1)).from_address(owner.ptr.value)
    a = np.frombuffer(buf, dtype=dtype, count=int(np.prod(shape))).reshape(shape)
    a_owner = a.base
    # keep the allocation alive as long as the array
    _keep[id(a_owner)] = owner
    import weakref
    weakref.finalize(a_owner, _keep.pop, id(a_owner), None)
    return a


_keep: dict[int, _Pinned] = {}


def grid_key(grid_input, grid_output, weights_input=None) -> int:
    """64-bit key of a grid pair's weights: a hash of the (perturbed)
    coordinates, their shapes and the input weighting."""
    import xxhash

    h = xxhash.xxh3_64()
    for a in (*grid_input, *grid_output) + ((weights_input,) if weights_input is not None else ()):
        a = np.ascontiguousarray(a, dtype=np.float64)
        h.update(np.asarray(a.shape, dtype=np.int64).tobytes())
        h.update(a.tobytes())
    k = h.intdigest()
    return k if k else 1


class Ticket:
    """A submitted request: its id, its result buffer and its arrival time.
    The host buffers the engine reads and writes are held by the server until
    the request is waited for or the server closes, so dropping a ticket
    early is safe."""

    __slots__ = ("id", "result", "arrival_ns")

    def __init__(self, tid, result, arrival_ns):
        self.id = tid
        self.result = result
        self.arrival_ns = arrival_ns


class Server:
    """A running serving loop.  ``submit`` is thread-safe and returns at
    once; ``wait`` blocks for one request."""

    def __init__(self, workers: int = 0, max_group: int = 0, reserve_bytes: int = 0,
                 cache_bytes: int = 0, edges_per_lane: int = 0, block: int = 0):
        L = _bind()
        cfg = rg_serve_config(workers, max_group, reserve_bytes, cache_bytes,
                              rg_tuning(edges_per_lane, block, 0, 0))
        self._h = ctypes.c_void_p()
        # host buffers of the requests not yet waited for, by ticket id
        self._pending: dict[int, list] = {}
        check(L.rg_serve_create(ctypes.byref(cfg), ctypes.byref(self._h)))

    def submit(self, values: np.ndarray, grid_input=None, grid_output=None, key: int = 0,
               shapes=None, weights_input=None, transpose: bool = False,
               result: np.ndarray | None = None, arrival_ns: int = 0) -> Ticket:
        """``values``: host (batch, cells) float64 or float32, C-contiguous.
        Give the (perturbed) grids, or a ``key`` whose weights are cached
        plus ``shapes=((nxi, nyi), (nxo, nyo))``; grids with a key are built
        once and cached.  ``result``: optional host output buffer."""
        L = _bind()
        values = np.asarray(values)
        if values.dtype not in (np.float64, np.float32) or not values.flags.c_contiguous:
            raise ValueError("values must be a C-contiguous float64 or float32 array")
        keep = [values]
        r = rg_serve_request()
        r.key = int(key)
        if grid_input is not None:
            xi, yi = (np.ascontiguousarray(a, dtype=np.float64) for a in grid_input)
            xo, yo = (np.ascontiguousarray(a, dtype=np.float64) for a in grid_output)
            keep += [xi, yi, xo, yo]
            r.xi, r.yi, r.xo, r.yo = (a.ctypes.data for a in (xi, yi, xo, yo))
            (r.nxi, r.nyi), (r.nxo, r.nyo) = xi.shape, xo.shape
            if weights_input is not None:
                wi = np.ascontiguousarray(weights_input, dtype=np.float64)
                keep.append(wi)
                r.wi = wi.ctypes.data
        elif shapes is not None:
            (r.nxi, r.nyi), (r.nxo, r.nyo) = shapes
        else:
            raise ValueError("give the grids or the shapes of cached weights")
        n_in = (r.nxi - 1) * (r.nyi - 1)
        n_out = (r.nxo - 1) * (r.nyo - 1)
        n_src, n_dst = (n_out, n_in) if transpose else (n_in, n_out)
        values2 = values.reshape(-1, n_src) if values.size else values.reshape(0, n_src)
        if result is None:
            result = np.empty((values2.shape[0], n_dst), dtype=values.dtype)
        if result.dtype != values.dtype or result.size != values2.shape[0] * n_dst \
                or not result.flags.c_contiguous:
            raise ValueError("result buffer has the wrong dtype, size or layout")
        keep.append(result)
        r.dtype = RG_F32 if values.dtype == np.float32 else RG_F64
        r.transpose = int(bool(transpose))
        r.values = values.ctypes.data
        r.result = result.ctypes.data
        r.batch = values2.shape[0]
        r.arrival_ns = int(arrival_ns)
        t = ctypes.c_uint64()
        check(L.rg_serve_submit(self._h, ctypes.byref(r), ctypes.byref(t)))
        self._pending[t.value] = keep
        return Ticket(t.value, result, r.arrival_ns)

    def wait(self, ticket: Ticket) -> tuple[np.ndarray, int]:
        """The request's result and its completion time (rg_clock_ns)."""
        L = _bind()
        done = ctypes.c_int64()
        status = L.rg_serve_wait(self._h, ticket.id, ctypes.byref(done))
        self._pending.pop(ticket.id, None)
        if status:
            raise RGError(status, L.rg_last_error().decode(errors="replace"))
        return ticket.result, int(done.value)

    def stats(self) -> dict:
        s = rg_serve_stats()
        check(_bind().rg_serve_stats_get(self._h, ctypes.byref(s)))
        return {name: getattr(s, name) for name, _ in rg_serve_stats._fields_}

    def close(self) -> None:
        """Finish every submitted request, stop the workers, free the cache;
        then release the buffers of requests nobody waited for."""
        if self._h and self._h.value:
            check(_bind().rg_serve_destroy(self._h))
            self._h = None
        self._pending.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
