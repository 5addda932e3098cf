"""
Weights on disk and on the wire, and a persistent store of built weights.

The reference hands weights around as ``(weights, shape_values_input,
shape_values_output)`` where ``weights`` is an object array of flat
``(indices_input, indices_output, values)`` triples, "a form that pickles
and memory-maps cleanly" (``regridding/_weights/_weights.py:31-36``,
``regridding/_regrid/_tests/test_regrid_from_weights.py:83-97``).  This
module gives that tuple one file layout, read back as memory-mapped views:

    offset 0   8 bytes  magic b"RGWGHT01"
    offset 8   8 bytes  little-endian uint64: length L of the JSON header
    offset 16  L bytes  JSON header, space-padded so that 16 + L is a
                        multiple of 64
    then three 64-byte aligned little-endian blocks (offsets in the header,
    relative to 16 + L) holding every element's triples back to back in the
    object array's C order:
               int64   indices_input   (sum of nnz)
               int64   indices_output  (sum of nnz)
               float64 values          (sum of nnz)

The header holds the format version, ``shape_orthogonal``,
``shape_values_input`` / ``shape_values_output``, each element's ``nnz``,
the byte offsets of the three blocks and, when the weights came from this
library's builders, the cells per element (``n_in``, ``n_out``) that let a
load import every element onto the GPU as one block-diagonal handle (the
same device layout the builders leave behind), so ``regrid_from_weights``
on loaded weights applies without re-uploading them.

:class:`WeightsStore` keeps weights in a directory keyed by
:func:`weights_key`, a 128-bit hash of everything that determines them:
the unperturbed coordinates (shape and bytes), the axes, ``weights_input``,
``method``, ``bounds``, ``perturb`` and the integer ``seed``.
"""

from __future__ import annotations

import json
import os
import tempfile
import threading
from typing import Sequence

import numpy as np

__all__ = [
    "save_weights",
    "load_weights",
    "weights_to_bytes",
    "weights_from_bytes",
    "weights_key",
    "WeightsStore",
]

MAGIC = b"RGWGHT01"
VERSION = 1
_ALIGN = 64


def _align(n: int) -> int:
    return (n + _ALIGN - 1) // _ALIGN * _ALIGN


def _element_cells(flat) -> tuple[int, int] | None:
    """(n_in, n_out) per element when every element is a registered device
    triple with the same block size, else None."""
    from . import _cache

    sizes = set()
    for ii, io, w in flat:
        if not (isinstance(ii, np.ndarray) and isinstance(io, np.ndarray)):
            return None
        hit = _cache.lookup_part(ii, io, np.asarray(w))
        if hit is None:
            return None
        handle, transposed, part = hit
        if part is None:
            bi, bo = handle.n_in, handle.n_out
        else:
            bi, bo = part[2], part[3]
        if transposed:
            bi, bo = bo, bi
        sizes.add((int(bi), int(bo)))
    return sizes.pop() if len(sizes) == 1 else None


def _layout(weights) -> tuple[bytes, list, int]:
    arr, shape_in, shape_out = weights
    arr = np.asarray(arr, dtype=object) if not isinstance(arr, np.ndarray) else arr
    flat = list(arr.reshape(-1))
    parts = []
    for element in flat:
        ii, io, w = element
        if getattr(w, "unit", None) is not None:
            raise TypeError("weights with units are not supported by the weights file format")
        ii = np.ascontiguousarray(ii, dtype="<i8").reshape(-1)
        io = np.ascontiguousarray(io, dtype="<i8").reshape(-1)
        w = np.ascontiguousarray(w, dtype="<f8").reshape(-1)
        if not (ii.size == io.size == w.size):
            raise ValueError("the three arrays of a weights element differ in length")
        parts.append((ii, io, w))
    nnz = [p[0].size for p in parts]
    total = int(sum(nnz))
    cells = _element_cells(flat) if flat else None
    block = _align(8 * total)
    header = {
        "version": VERSION,
        "shape_orthogonal": list(arr.shape),
        "shape_values_input": [int(v) for v in shape_in],
        "shape_values_output": [int(v) for v in shape_out],
        "nnz": nnz,
        "cells": list(cells) if cells else None,
        # block offsets relative to the end of the (padded) header
        "offsets": [0, block, 2 * block],
    }
    text = json.dumps(header, separators=(",", ":")).encode()
    text = text + b" " * (_align(16 + len(text)) - 16 - len(text))
    return MAGIC + np.uint64(len(text)).astype("<u8").tobytes() + text, parts, block


def _write(f, weights) -> None:
    head, parts, block = _layout(weights)
    f.write(head)
    for k in range(3):
        written = 0
        for p in parts:
            f.write(p[k].tobytes())
            written += p[k].nbytes
        f.write(b"\0" * (block - written))


def weights_to_bytes(weights) -> bytes:
    """The file layout as one ``bytes`` object (for a socket or a queue)."""
    import io

    buf = io.BytesIO()
    _write(buf, weights)
    return buf.getvalue()


def save_weights(file: str | os.PathLike, weights) -> None:
    """Write ``weights`` (the tuple :func:`regridding_amd.weights` returns)
    to ``file``.  The file is written under a temporary name and renamed,
    so a reader never sees a partial file."""
    file = os.fspath(file)
    d = os.path.dirname(os.path.abspath(file))
    fd, tmp = tempfile.mkstemp(prefix=".rgw-", dir=d)
    try:
        with os.fdopen(fd, "wb") as f:
            _write(f, weights)
        os.replace(tmp, file)
    except BaseException:
        if os.path.exists(tmp):
            os.unlink(tmp)
        raise


def _parse(buf) -> tuple[dict, np.ndarray, np.ndarray, np.ndarray]:
    mv = memoryview(buf)
    if len(mv) < 16 or bytes(mv[:8]) != MAGIC:
        raise ValueError("not a regridding weights file (bad magic)")
    hlen = int(np.frombuffer(mv[8:16], dtype="<u8")[0])
    if 16 + hlen > len(mv):
        raise ValueError("truncated weights file (header)")
    try:
        header = json.loads(bytes(mv[16:16 + hlen]).decode())
    except (UnicodeDecodeError, json.JSONDecodeError) as e:
        raise ValueError(f"corrupt weights file header: {e}") from None
    if header.get("version") != VERSION:
        raise ValueError(f"unsupported weights file version {header.get('version')!r}")
    nnz = [int(v) for v in header["nnz"]]
    if any(v < 0 for v in nnz) or len(nnz) != int(np.prod(header["shape_orthogonal"])):
        raise ValueError("corrupt weights file header (element counts)")
    total = sum(nnz)
    data0 = 16 + hlen
    o = [data0 + int(v) for v in header["offsets"]]
    if data0 % _ALIGN or any(v % _ALIGN for v in o) or o[2] + 8 * total > len(mv) or \
            o[1] < o[0] + 8 * total or o[2] < o[1] + 8 * total:
        raise ValueError("truncated or corrupt weights file (data)")
    ii = np.frombuffer(mv, dtype="<i8", count=total, offset=o[0])
    io = np.frombuffer(mv, dtype="<i8", count=total, offset=o[1])
    w = np.frombuffer(mv, dtype="<f8", count=total, offset=o[2])
    return header, ii, io, w


def _assemble(header, ii, io, w, device: bool):
    nnz = header["nnz"]
    starts = np.concatenate([[0], np.cumsum(nnz, dtype=np.int64)]).astype(np.int64)
    parts = [(ii[s:e], io[s:e], w[s:e]) for s, e in zip(starts[:-1], starts[1:])]
    for p in parts:
        for a in p:
            a.flags.writeable = False
    shape_orthogonal = tuple(header["shape_orthogonal"])
    cells = header.get("cells")
    if device and cells and len(parts) == 1:
        # one element: a whole handle (found again by every broadcast use)
        from . import _cache
        from ._regrid import import_weights

        handle = import_weights(*parts[0], int(cells[0]), int(cells[1]))
        _cache.register(*parts[0], handle)
        result = np.empty(1, dtype=object)
        result[0] = parts[0]
        result = result.reshape(shape_orthogonal)
    elif device and cells and parts:
        from ._line import _register_parts, block_handle

        handle = block_handle(parts, int(cells[0]), int(cells[1]))
        result = _register_parts(parts, handle, int(cells[0]), int(cells[1]), shape_orthogonal)
    else:
        result = np.empty(len(parts), dtype=object)
        for k, p in enumerate(parts):
            result[k] = p
        result = result.reshape(shape_orthogonal)
    return (result, tuple(header["shape_values_input"]), tuple(header["shape_values_output"]))


def weights_from_bytes(buf, device: bool = False):
    """Inverse of :func:`weights_to_bytes`; the arrays are read-only views
    of ``buf``.  ``device=True`` also imports them onto the GPU (see
    :func:`load_weights`)."""
    return _assemble(*_parse(buf), device=device)


def load_weights(file: str | os.PathLike, mmap: bool = True, device: bool = False):
    """
    Read weights written by :func:`save_weights`: the reference's tuple
    ``(weights, shape_values_input, shape_values_output)`` whose arrays are
    read-only views of the memory-mapped file (``mmap=False`` reads it into
    memory instead).

    ``device=True`` imports every element onto the GPU as one block-diagonal
    handle (when the file records the cells per element), so that
    ``regrid_from_weights`` on the returned arrays applies without uploading
    them again.
    """
    file = os.fspath(file)
    if mmap:
        if os.path.getsize(file) == 0:
            raise ValueError("not a regridding weights file (empty)")
        buf = np.memmap(file, dtype=np.uint8, mode="r")
    else:
        with open(file, "rb") as f:
            buf = f.read()
    return _assemble(*_parse(buf), device=device)


def _seed_is_deterministic(seed) -> bool:
    return isinstance(seed, (int, np.integer)) and not isinstance(seed, bool)


def weights_key(
    coordinates_input,
    coordinates_output,
    axis_input: None | int | Sequence[int] = None,
    axis_output: None | int | Sequence[int] = None,
    weights_input: None | np.ndarray = None,
    method: str = "multilinear",
    bounds: str = "extrapolate",
    perturb: None | bool = None,
    seed=42,
) -> str | None:
    """
    Hex key (xxh3-128) of the weights that :func:`regridding_amd.weights`
    returns for these arguments, or None when the result is not a function
    of them alone (``seed`` None or a ``Generator`` while the output grid is
    perturbed: those draw from fresh or caller-held random state).
    """
    import xxhash

    from . import _util

    if method == "conservative":
        eff_perturb = _util.resolve_perturb(perturb, coordinates_input)
    else:
        eff_perturb = bool(perturb)
    if eff_perturb and not _seed_is_deterministic(seed):
        return None
    if getattr(weights_input, "unit", None) is not None:
        return None
    h = xxhash.xxh3_128()

    def put(tag: str, value) -> None:
        h.update(tag.encode() + b"\0" + json.dumps(value).encode() + b"\0")

    def put_array(tag: str, a) -> None:
        a = np.asarray(a)
        put(tag, [str(a.dtype), list(a.shape)])
        h.update(np.ascontiguousarray(a).tobytes())

    put("format", VERSION)
    put("method", method)
    put("bounds", bounds)
    put("perturb", [perturb if perturb is None else bool(perturb), eff_perturb])
    put("seed", int(seed) if eff_perturb else None)
    for name, ax in (("axis_input", axis_input), ("axis_output", axis_output)):
        put(name, None if ax is None else np.atleast_1d(ax).astype(int).tolist())
    for name, coords in (("input", coordinates_input), ("output", coordinates_output)):
        if isinstance(coords, np.ndarray):
            put(name, "array")
            put_array(name, coords)
        else:
            put(name, len(coords))
            for c in coords:
                put_array(name, c)
    if weights_input is None:
        put("weights_input", None)
    else:
        put_array("weights_input", weights_input)
    return h.hexdigest()


class WeightsStore:
    """
    A directory of weights files keyed by :func:`weights_key`.

    :meth:`weights` has the signature of :func:`regridding_amd.weights`; a
    hit returns the stored weights (memory-mapped, imported onto the GPU),
    a miss builds them and stores them.  ``max_bytes`` bounds the directory:
    the least recently used files are removed after a store.  Safe for
    several processes sharing one directory (files are renamed into place).
    """

    def __init__(self, directory: str | os.PathLike, max_bytes: int | None = None,
                 device: bool = True):
        self.directory = os.fspath(directory)
        os.makedirs(self.directory, exist_ok=True)
        self.max_bytes = max_bytes
        self.device = device
        self.hits = 0
        self.misses = 0
        self._lock = threading.Lock()

    def path(self, key: str) -> str:
        return os.path.join(self.directory, key + ".rgw")

    def get(self, key: str):
        p = self.path(key)
        try:
            result = load_weights(p, device=self.device)
        except (FileNotFoundError, ValueError):
            return None
        try:
            os.utime(p)
        except OSError:
            pass
        return result

    def put(self, key: str, weights) -> None:
        save_weights(self.path(key), weights)
        if self.max_bytes is not None:
            self._evict()

    def _evict(self) -> None:
        files = []
        for name in os.listdir(self.directory):
            if not name.endswith(".rgw"):
                continue
            p = os.path.join(self.directory, name)
            try:
                st = os.stat(p)
            except FileNotFoundError:
                continue
            files.append((st.st_mtime, st.st_size, p))
        files.sort()
        total = sum(f[1] for f in files)
        for _, size, p in files[:-1]:
            if total <= self.max_bytes:
                break
            try:
                os.unlink(p)
                total -= size
            except FileNotFoundError:
                pass

    def weights(self, coordinates_input, coordinates_output, axis_input=None, axis_output=None,
                weights_input=None, method="multilinear", bounds="extrapolate", perturb=None,
                seed=42):
        args = dict(axis_input=axis_input, axis_output=axis_output,
                    weights_input=weights_input, method=method, bounds=bounds,
                    perturb=perturb, seed=seed)
        key = weights_key(coordinates_input, coordinates_output, **args)
        if key is not None:
            hit = self.get(key)
            if hit is not None:
                with self._lock:
                    self.hits += 1
                return hit
        result = _build(coordinates_input, coordinates_output, **args)
        with self._lock:
            self.misses += 1
        if key is not None:
            self.put(key, result)
        return result


def _build(coordinates_input, coordinates_output, **kwargs):
    from ._weights import weights

    return weights(coordinates_input, coordinates_output, **kwargs)
