"""
Device residency of weights returned to the host.

``weights()`` must return host arrays (the reference's flat
``(indices_input, indices_output, values)`` triples,
``regridding/_weights/_weights.py:31-36``), but the device already holds
them as CSR after the build.  This registry remembers, for each returned
triple, the device handle it came from, so ``regrid_from_weights`` on those
very arrays (or on their ``transpose_weights`` swap) applies the resident
CSR instead of uploading and regrouping the triples again.

Entries are keyed by the identity of the three arrays and die with them.
A triple may also be registered as part ``d`` of ``D`` of a block-diagonal
handle (the 1-D builders put all elements of one call in one handle).
The arrays stay writable, as the reference returns them
(``_weights_arrays.py:68-72``): each entry records an xxh3 digest of the
three arrays' bytes, and a hit counts only while the arrays still hash to
it, so a caller who edits the weights in place gets the edited weights
applied (one pass over the arrays, about 7 GB/s on one core, per lookup).
Arrays the registry has not seen (unpickled, hand-made, modified) go
through the import path.
"""

from __future__ import annotations

import threading
import weakref

import numpy as np
import xxhash

from ._lib import Weights

__all__ = ["register", "lookup"]

_lock = threading.Lock()
# id(values) -> (weakref(indices_input), weakref(indices_output), Weights, part, digest)
_entries: dict[int, tuple] = {}


def _drop(key: int) -> None:
    with _lock:
        _entries.pop(key, None)


def _digest(a: np.ndarray, b: np.ndarray, values: np.ndarray) -> tuple | None:
    """Content digest of a triple (input and output roles told apart by
    their order), or None when an array is not C-contiguous."""
    if not (a.flags.c_contiguous and b.flags.c_contiguous and values.flags.c_contiguous):
        return None
    return tuple((x.dtype.str, x.shape, xxhash.xxh3_64_intdigest(x)) for x in (a, b, values))


def register(indices_input: np.ndarray, indices_output: np.ndarray, values: np.ndarray,
             handle: Weights, part: tuple[int, int, int, int] | None = None) -> None:
    """``part``: ``(d, D, n_in, n_out)`` when ``handle`` is block-diagonal
    and this triple is its block d (n_in x n_out cells per block)."""
    digest = _digest(indices_input, indices_output, values)
    if digest is None:
        return
    key = id(values)
    with _lock:
        _entries[key] = (weakref.ref(indices_input), weakref.ref(indices_output), handle, part,
                         digest)
    weakref.finalize(values, _drop, key)


def lookup(a: np.ndarray, b: np.ndarray, values: np.ndarray) -> tuple[Weights, bool] | None:
    """
    ``(handle, transposed)`` if ``(a, b, values)`` is a registered whole
    triple (``transposed=False``) or its index swap (``transposed=True``).
    """
    hit = lookup_part(a, b, values)
    if hit is None or hit[2] is not None:
        return None
    return hit[0], hit[1]


def lookup_part(a: np.ndarray, b: np.ndarray, values: np.ndarray):
    """``(handle, transposed, part)`` for a registered triple or its swap
    (``part`` None for a whole handle), else None."""
    with _lock:
        e = _entries.get(id(values))
    if e is None:
        return None
    ri, ro, handle, part, digest = e
    ii, io = ri(), ro()
    if ii is None or io is None:
        return None
    if a is ii and b is io:
        transposed = False
    elif a is io and b is ii:
        transposed = True
    else:
        return None
    if _digest(ii, io, values) != digest:
        # edited in place since the build: the resident CSR is stale
        _drop(id(values))
        return None
    return handle, transposed, part
