"""
ctypes binding of ``librg_hip.so`` (the HIP/gfx950 library, C ABI in
``include/regridding_amd.h``) and small device-memory helpers.

There is no CPU fallback: if the library is missing or no GPU is visible,
every entry point that needs it raises ``RuntimeError``.
"""

from __future__ import annotations

import ctypes
import importlib.util
import os
import sys
import threading
import weakref

import numpy as np

__all__ = [
    "RGError",
    "lib",
    "check",
    "DeviceArray",
    "Weights",
    "EventTimer",
    "LIB_PATH",
]

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "librg_hip.so")

_i64 = ctypes.c_int64
_vp = ctypes.c_void_p
_pi64 = ctypes.POINTER(ctypes.c_int64)


class RGError(RuntimeError):
    """A failure reported by the HIP library (status code in ``.code``)."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[rg status {code}] {message}")
        self.code = code


class rg_tuning(ctypes.Structure):
    _fields_ = [
        ("edges_per_lane", ctypes.c_int32),
        ("block", ctypes.c_int32),
        ("flags", ctypes.c_int32),
        ("stage_per_edge", ctypes.c_int32),
    ]


class rg_pair_desc(ctypes.Structure):
    _fields_ = [
        ("xi", ctypes.c_void_p),
        ("yi", ctypes.c_void_p),
        ("nxi", ctypes.c_int64),
        ("nyi", ctypes.c_int64),
        ("xo", ctypes.c_void_p),
        ("yo", ctypes.c_void_p),
        ("nxo", ctypes.c_int64),
        ("nyo", ctypes.c_int64),
        ("wi", ctypes.c_void_p),
    ]


class rg_apply_job(ctypes.Structure):
    _fields_ = [
        ("weights", ctypes.c_void_p),
        ("transpose", ctypes.c_int32),
        ("values_input", ctypes.c_void_p),
        ("values_output", ctypes.c_void_p),
        ("batch", ctypes.c_int64),
    ]


class rg_serve_config(ctypes.Structure):
    _fields_ = [
        ("workers", ctypes.c_int32),
        ("max_group", ctypes.c_int32),
        ("reserve_bytes", ctypes.c_int64),
        ("cache_bytes", ctypes.c_int64),
        ("tuning", rg_tuning),
    ]


class rg_serve_request(ctypes.Structure):
    _fields_ = [
        ("key", ctypes.c_uint64),
        ("xi", ctypes.c_void_p),
        ("yi", ctypes.c_void_p),
        ("nxi", ctypes.c_int64),
        ("nyi", ctypes.c_int64),
        ("xo", ctypes.c_void_p),
        ("yo", ctypes.c_void_p),
        ("nxo", ctypes.c_int64),
        ("nyo", ctypes.c_int64),
        ("wi", ctypes.c_void_p),
        ("dtype", ctypes.c_int32),
        ("transpose", ctypes.c_int32),
        ("values", ctypes.c_void_p),
        ("result", ctypes.c_void_p),
        ("batch", ctypes.c_int64),
        ("arrival_ns", ctypes.c_int64),
    ]


class rg_serve_stats(ctypes.Structure):
    _fields_ = [(name, ctypes.c_int64) for name in (
        "submitted", "completed", "failed", "cache_hits", "cache_misses", "builds",
        "groups", "max_group_seen", "peak_admitted_bytes", "cache_entries", "cache_bytes")]


class rg_build_stats(ctypes.Structure):
    _fields_ = [
        ("n_triples", ctypes.c_int64),
        ("n_chunks", ctypes.c_int64),
        ("n_fixups", ctypes.c_int64),
        ("n_subsegments", ctypes.c_int64),
        ("n_search_anomalies", ctypes.c_int64),
        ("n_restaged", ctypes.c_int64),
        ("n_long_cells", ctypes.c_int64),
        ("ms_prepare", ctypes.c_double),
        ("ms_start", ctypes.c_double),
        ("ms_count", ctypes.c_double),
        ("ms_fixup", ctypes.c_double),
        ("ms_emit", ctypes.c_double),
        ("ms_reduce", ctypes.c_double),
        ("ms_transpose", ctypes.c_double),
        ("ms_total", ctypes.c_double),
    ]


RG_TUNE_TIMINGS = 1
RG_FIND_BRUTE = 0
RG_FIND_SEARCHSORTED = 1

# name -> (restype, argtypes); every function of include/regridding_amd.h
SIGNATURES = {
    "rg_last_error": (ctypes.c_char_p, []),
    "rg_version": (ctypes.c_int, []),
    "rg_init": (ctypes.c_int, [ctypes.c_int]),
    "rg_device_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
    "rg_mem_info": (ctypes.c_int, [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]),
    "rg_trim": (None, []),
    "rg_malloc": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_size_t]),
    "rg_free": (ctypes.c_int, [_vp]),
    "rg_memcpy_htod": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t, _vp]),
    "rg_memcpy_dtoh": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t, _vp]),
    "rg_memset": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_size_t, _vp]),
    "rg_stream_create": (ctypes.c_int, [ctypes.POINTER(_vp)]),
    "rg_stream_destroy": (ctypes.c_int, [_vp]),
    "rg_stream_sync": (ctypes.c_int, [_vp]),
    "rg_device_sync": (ctypes.c_int, []),
    "rg_event_create": (ctypes.c_int, [ctypes.POINTER(_vp)]),
    "rg_event_destroy": (ctypes.c_int, [_vp]),
    "rg_event_record": (ctypes.c_int, [_vp, _vp]),
    "rg_event_elapsed_ms": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(ctypes.c_float)]),
    "rg_copy_f64": (ctypes.c_int, [_vp, _vp, _i64, _vp]),
    "rg_cell_area_2d": (ctypes.c_int, [_vp, _vp, _i64, _i64, _vp, _vp]),
    "rg_build_conservative_2d": (
        ctypes.c_int,
        [_vp, _vp, _i64, _i64, _vp, _vp, _i64, _i64, _vp, ctypes.POINTER(rg_tuning), ctypes.POINTER(_vp), _vp],
    ),
    "rg_build_conservative_2d_batched": (
        ctypes.c_int,
        [ctypes.POINTER(rg_pair_desc), _i64, ctypes.POINTER(rg_tuning), ctypes.POINTER(_vp), _vp],
    ),
    "rg_apply_batched_f64": (ctypes.c_int, [ctypes.POINTER(rg_apply_job), _i64, _vp]),
    "rg_apply_batched_f32": (ctypes.c_int, [ctypes.POINTER(rg_apply_job), _i64, _vp]),
    "rg_tune_build": (
        ctypes.c_int,
        [ctypes.POINTER(rg_pair_desc), _i64, _vp, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int,
         ctypes.POINTER(rg_tuning), ctypes.POINTER(rg_tuning), ctypes.POINTER(ctypes.c_double),
         _vp, _vp],
    ),
    "rg_host_alloc": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_size_t]),
    "rg_host_free": (ctypes.c_int, [_vp]),
    "rg_clock_ns": (ctypes.c_int64, []),
    "rg_serve_create": (ctypes.c_int, [ctypes.POINTER(rg_serve_config), ctypes.POINTER(_vp)]),
    "rg_serve_submit": (ctypes.c_int, [_vp, ctypes.POINTER(rg_serve_request),
                                       ctypes.POINTER(ctypes.c_uint64)]),
    "rg_serve_wait": (ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int64)]),
    "rg_serve_stats_get": (ctypes.c_int, [_vp, ctypes.POINTER(rg_serve_stats)]),
    "rg_serve_destroy": (ctypes.c_int, [_vp]),
    "rg_conservative_1d_count": (ctypes.c_int, [_vp, _i64, _vp, _i64, _i64, _vp, _vp, _pi64, _vp]),
    "rg_conservative_1d_emit": (ctypes.c_int, [_vp, _i64, _vp, _i64, _i64, _vp, _vp, _vp, _vp,
                                               _vp, _vp, _vp]),
    "rg_cell_length_1d": (ctypes.c_int, [_vp, _i64, _i64, _vp, _vp]),
    "rg_graph_create": (ctypes.c_int, [_vp, _i64, _vp, _vp, ctypes.c_double, _vp, _vp]),
    "rg_graph_launch": (ctypes.c_int, [_vp, _vp]),
    "rg_graph_status": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int32), _vp]),
    "rg_graph_weights": (ctypes.c_int, [_vp, _i64, _vp, _vp]),
    "rg_graph_capacity": (ctypes.c_int, [_vp, _pi64, _pi64, _pi64, _pi64]),
    "rg_graph_destroy": (ctypes.c_int, [_vp]),
    "rg_fill_gauss_seidel_2d": (ctypes.c_int, [_vp, _vp, _i64, _i64, _i64, _i64, _vp]),
    "rg_interp_linear_1d": (ctypes.c_int, [_vp, _i64, _vp, _i64, _i64, _vp, _vp]),
    "rg_interp_linear_2d": (ctypes.c_int, [_vp, _i64, _i64, _vp, _vp, _i64, _i64, _vp, _vp]),
    "rg_multilinear_1d": (ctypes.c_int, [_vp, _i64, _vp, _i64, _i64, _vp, _vp, ctypes.c_int32,
                                         _vp, _vp, _vp, _pi64, _vp]),
    "rg_weights_info": (ctypes.c_int, [_vp, _pi64, _pi64, _pi64]),
    "rg_weights_stats": (ctypes.c_int, [_vp, ctypes.POINTER(rg_build_stats)]),
    "rg_weights_export": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "rg_weights_import": (ctypes.c_int, [_vp, _vp, _vp, _i64, _i64, _i64, ctypes.POINTER(_vp), _vp]),
    "rg_weights_free": (ctypes.c_int, [_vp]),
    "rg_weights_set_shape": (ctypes.c_int, [_vp, _i64, _i64, _i64, _i64]),
    "rg_set_apply_tuning": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "rg_apply_f64": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _i64, _vp]),
    "rg_apply_f32": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _i64, _vp]),
    "rg_find_cells_2d": (ctypes.c_int, [_vp, _vp, _i64, _i64, _vp, _vp, _i64, _vp, _vp]),
    "rg_find_cells_2d_noalloc": (ctypes.c_int, [_vp, _vp, _i64, _i64, _vp, _vp, _i64, _vp, _vp]),
    "rg_find_indices_1d": (ctypes.c_int, [_vp, _i64, _vp, _i64, _i64, ctypes.c_int, _i64, _vp, _vp]),
    "rg_inject_failure": (ctypes.c_int, [ctypes.c_char_p]),
}

_lock = threading.Lock()
_lib = None


def _load_torch_runtime_first() -> None:
    """PyTorch-ROCm carries its own HIP runtime under the soname this
    library links (libamdhip64.so.7), and a process can hold only one: the
    first loaded serves both (a torch imported after the ROCm one finds no
    GPU).  When torch is installed, its runtime library is loaded first,
    globally, from torch's own lib folder, without importing torch, so this
    library and a torch imported later share it; if that load fails, torch
    itself is imported."""
    if os.environ.get("RG_NO_TORCH") or "torch" in sys.modules:
        return
    spec = importlib.util.find_spec("torch")
    if spec is None or not spec.submodule_search_locations:
        return
    runtime = os.path.join(list(spec.submodule_search_locations)[0], "lib", "libamdhip64.so")
    try:
        ctypes.CDLL(runtime, mode=ctypes.RTLD_GLOBAL)
        return
    except OSError:
        pass
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load the shared library and declare its signatures (no GPU call)."""
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: the HIP extension of regridding_amd is not "
            "built (run __graft_entry__.build() or make -C regridding_amd/csrc)"
        )
    _load_torch_runtime_first()
    L = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def lib() -> ctypes.CDLL:
    """The initialised library (device 0 selected on first use)."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                L = load_library()
                device = int(os.environ.get("RG_DEVICE", "0"))
                st = L.rg_init(device)
                if st != 0:
                    raise RGError(st, L.rg_last_error().decode() + " (is a GPU visible?)")
                _lib = L
    return _lib


def check(status: int) -> None:
    if status != 0:
        raise RGError(status, lib().rg_last_error().decode())


def _finalize_ptr(ptr: int) -> None:
    if _lib is not None and ptr:
        _lib.rg_free(ptr)


class rg_graph_job(ctypes.Structure):
    _fields_ = [
        ("values_input", ctypes.c_void_p),
        ("values_output", ctypes.c_void_p),
        ("batch", ctypes.c_int64),
    ]


class DeviceArray:
    """A block of device memory owned by the library's caching allocator."""

    def __init__(self, nbytes: int, dtype=np.float64, shape: tuple = ()):
        self.nbytes = int(nbytes)
        self.dtype = np.dtype(dtype)
        self.shape = tuple(shape)
        p = _vp()
        check(lib().rg_malloc(ctypes.byref(p), max(self.nbytes, 1)))
        self.ptr = p.value or 0
        self._fin = weakref.finalize(self, _finalize_ptr, self.ptr)

    @classmethod
    def empty(cls, shape, dtype=np.float64) -> "DeviceArray":
        shape = tuple(int(s) for s in np.atleast_1d(shape)) if np.ndim(shape) else (int(shape),)
        n = int(np.prod(shape)) * np.dtype(dtype).itemsize
        return cls(n, dtype, shape)

    @classmethod
    def from_host(cls, a: np.ndarray, stream: int = 0) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes, a.dtype, a.shape)
        if a.nbytes:
            check(lib().rg_memcpy_htod(d.ptr, a.ctypes.data, a.nbytes, stream))
        return d

    def copy_from_host(self, a: np.ndarray, stream: int = 0) -> None:
        """Overwrite the block with `a` (same byte size)."""
        a = np.ascontiguousarray(a)
        if a.nbytes != self.nbytes:
            raise ValueError(f"{a.nbytes} bytes do not fit a block of {self.nbytes}")
        if a.nbytes:
            check(lib().rg_memcpy_htod(self.ptr, a.ctypes.data, a.nbytes, stream))

    def to_host(self, stream: int = 0, out: np.ndarray | None = None) -> np.ndarray:
        if out is None:
            out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            check(lib().rg_memcpy_dtoh(out.ctypes.data, self.ptr, self.nbytes, stream))
        return out

    def free(self) -> None:
        self._fin()


class EventTimer:
    """A pair of HIP events recorded on one stream."""

    def __init__(self, stream: int = 0):
        self.stream = stream
        a, b = _vp(), _vp()
        check(lib().rg_event_create(ctypes.byref(a)))
        check(lib().rg_event_create(ctypes.byref(b)))
        self.a, self.b = a.value, b.value

    def start(self) -> None:
        check(lib().rg_event_record(self.a, self.stream))

    def stop(self) -> float:
        """Record the stop event, wait for it, return milliseconds."""
        check(lib().rg_event_record(self.b, self.stream))
        ms = ctypes.c_float()
        check(lib().rg_event_elapsed_ms(self.a, self.b, ctypes.byref(ms)))
        return float(ms.value)

    def __del__(self):
        if _lib is not None:
            _lib.rg_event_destroy(self.a)
            _lib.rg_event_destroy(self.b)


def _finalize_weights(ptr: int) -> None:
    if _lib is not None and ptr:
        _lib.rg_weights_free(ptr)


class Weights:
    """Owner of an ``rg_weights_t`` handle: device CSR by input and by output."""

    def __init__(self, ptr: int):
        self.ptr = ptr
        n_in, n_out, nnz = _i64(), _i64(), _i64()
        check(lib().rg_weights_info(ptr, ctypes.byref(n_in), ctypes.byref(n_out), ctypes.byref(nnz)))
        self.n_in, self.n_out, self.nnz = n_in.value, n_out.value, nnz.value
        self._fin = weakref.finalize(self, _finalize_weights, ptr)

    def stats(self) -> dict:
        s = rg_build_stats()
        check(lib().rg_weights_stats(self.ptr, ctypes.byref(s)))
        return {name: getattr(s, name) for name, _ in rg_build_stats._fields_}

    def export(self, stream: int = 0) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """The coalesced triples sorted by (input, output), on the host."""
        n = self.nnz
        di = DeviceArray.empty(n, np.int64)
        do = DeviceArray.empty(n, np.int64)
        dw = DeviceArray.empty(n, np.float64)
        check(lib().rg_weights_export(self.ptr, di.ptr, do.ptr, dw.ptr, stream))
        return di.to_host(stream), do.to_host(stream), dw.to_host(stream)

    def free(self) -> None:
        self._fin()
