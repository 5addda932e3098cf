"""
Autotuner of the conservative build's launch shape.

The build's result does not depend on its tuning (edges walked per lane,
threads per block); its speed does, through the balance of chunk count,
fix-up work and occupancy, which shifts with the grid sizes in a batch.
``tune`` times every candidate with ``rg_tune_build`` (HIP events, median of
``reps`` runs after one untimed run) and caches the winner per batch
signature -- the rounded log2 of the total sweep edges and of the mean row
length -- in process, and in a JSON file when ``RG_TUNE_CACHE`` names one.
"""

from __future__ import annotations

import ctypes
import json
import math
import os
import threading

import numpy as np

from ._conservative import PackedPairs, pack_pairs
from ._lib import check, lib, rg_tuning

__all__ = ["EDGES_PER_LANE", "BLOCKS", "signature", "tune", "cached"]

EDGES_PER_LANE = (1, 2, 4, 8, 10, 12, 16, 32)
BLOCKS = (64, 128, 256, 512)

_lock = threading.Lock()
_cache: dict[str, dict] = {}
_loaded = False


def signature(packed: PackedPairs) -> str:
    """Size class of a batch: rounded log2 of its sweep edges and of the
    mean sweep-row length."""
    edges, rows = 0, 0
    for k in range(packed.n):
        d = packed.desc[k]
        for nx, ny in ((d.nxi, d.nyi), (d.nxo, d.nyo)):
            edges += nx * (ny - 1) + ny * (nx - 1)
            rows += nx + ny
    return f"e{round(math.log2(max(edges, 1)))}-r{round(math.log2(max(edges / max(rows, 1), 1)))}"


def _path():
    return os.environ.get("RG_TUNE_CACHE")


def _load():
    global _loaded
    if _loaded:
        return
    _loaded = True
    p = _path()
    if p and os.path.exists(p):
        try:
            with open(p) as f:
                data = json.load(f)
            if isinstance(data, dict):
                _cache.update({k: v for k, v in data.items() if isinstance(v, dict)})
        except (OSError, ValueError):
            pass


def _store():
    p = _path()
    if not p:
        return
    tmp = p + ".tmp"
    with open(tmp, "w") as f:
        json.dump(_cache, f, indent=1, sort_keys=True)
    os.replace(tmp, p)


def cached(packed: PackedPairs) -> dict | None:
    """The cached choice for this batch's size class, or None."""
    with _lock:
        _load()
        return _cache.get(signature(packed))


def tune(pairs, reps: int = 2, edges_per_lane=EDGES_PER_LANE, blocks=BLOCKS,
         stream: int = 0, force: bool = False) -> dict:
    """Pick edges_per_lane and block for building ``pairs`` (a PackedPairs
    or a sequence accepted by pack_pairs).  Returns ``{"edges_per_lane",
    "block", "ms", "table"}``; a cached choice for the size class is returned
    without timing unless ``force``."""
    packed = pairs if isinstance(pairs, PackedPairs) else pack_pairs(pairs)
    key = signature(packed)
    with _lock:
        _load()
        if not force and key in _cache:
            return dict(_cache[key])
    epl = np.asarray(edges_per_lane, dtype=np.int32)
    blk = np.asarray(blocks, dtype=np.int32)
    times = np.zeros((epl.size, blk.size), dtype=np.float64)
    best = rg_tuning()
    best_ms = ctypes.c_double()
    check(lib().rg_tune_build(packed.desc, packed.n, epl.ctypes.data, int(epl.size),
                              blk.ctypes.data, int(blk.size), int(reps), None,
                              ctypes.byref(best), ctypes.byref(best_ms),
                              times.ctypes.data, stream))
    choice = {
        "edges_per_lane": int(best.edges_per_lane),
        "block": int(best.block),
        "ms": round(float(best_ms.value), 4),
        "table": {f"{int(e)}x{int(b)}": round(float(times[i, j]), 4)
                  for i, e in enumerate(epl) for j, b in enumerate(blk)},
    }
    with _lock:
        _cache[key] = choice
        _store()
    return dict(choice)
