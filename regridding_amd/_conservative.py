"""
Device build of conservative 2-D weights for one orthogonal element.

Replaces the compiled part of ``regridding._weights._weights_conservative``
(``_weights_conservative.py:122-139`` calling ``weights_conservative_2d``,
``_weights_conservative_2d.py:80-126``, then ``_weights_arrays._coalesce``):
the coordinates (already normalised and perturbed on the host) are copied to
the device, ``rg_build_conservative_2d`` sweeps and coalesces, and the
handle keeps both CSR layouts resident for the apply.
"""

from __future__ import annotations

import ctypes

import numpy as np

from ._lib import DeviceArray, Weights, check, lib, rg_pair_desc, rg_tuning, RG_TUNE_TIMINGS

__all__ = ["build_2d", "build_2d_device", "build_2d_batched", "pack_pairs"]


def build_2d_device(
    xi: DeviceArray, yi: DeviceArray, shape_input: tuple[int, int],
    xo: DeviceArray, yo: DeviceArray, shape_output: tuple[int, int],
    wi: DeviceArray | None = None,
    edges_per_lane: int = 0, block: int = 0, timings: bool = False,
    stage_per_edge: int = 0, stream: int = 0, flags: int = 0,
) -> Weights:
    """Build from device-resident vertex arrays (row-major float64).
    ``flags`` adds ``RG_TUNE_*`` bits (test hooks) to the tuning."""
    t = rg_tuning(int(edges_per_lane), int(block),
                  (RG_TUNE_TIMINGS if timings else 0) | int(flags), int(stage_per_edge))
    h = ctypes.c_void_p()
    check(
        lib().rg_build_conservative_2d(
            xi.ptr, yi.ptr, int(shape_input[0]), int(shape_input[1]),
            xo.ptr, yo.ptr, int(shape_output[0]), int(shape_output[1]),
            wi.ptr if wi is not None else None, ctypes.byref(t), ctypes.byref(h), stream,
        )
    )
    return Weights(h.value)


def build_2d(
    grid_input: tuple[np.ndarray, np.ndarray],
    grid_output: tuple[np.ndarray, np.ndarray],
    weights_input: np.ndarray | None = None,
    **tuning,
) -> Weights:
    """Build from host vertex arrays of shape ``(nx, ny)``."""
    xi = np.ascontiguousarray(grid_input[0], dtype=np.float64)
    yi = np.ascontiguousarray(grid_input[1], dtype=np.float64)
    xo = np.ascontiguousarray(grid_output[0], dtype=np.float64)
    yo = np.ascontiguousarray(grid_output[1], dtype=np.float64)
    if xi.ndim != 2 or xo.ndim != 2 or xi.shape != yi.shape or xo.shape != yo.shape:
        raise ValueError("2-D grids need two coordinate arrays of one 2-D shape")
    wi = None
    if weights_input is not None:
        w = np.ascontiguousarray(weights_input, dtype=np.float64)
        if w.shape != (xi.shape[0] - 1, xi.shape[1] - 1):
            raise ValueError(
                f"weights_input has shape {w.shape}, expected the input cell "
                f"shape {(xi.shape[0] - 1, xi.shape[1] - 1)}"
            )
        wi = DeviceArray.from_host(w)
    return build_2d_device(
        DeviceArray.from_host(xi), DeviceArray.from_host(yi), xi.shape,
        DeviceArray.from_host(xo), DeviceArray.from_host(yo), xo.shape,
        wi, **tuning,
    )


class PackedPairs:
    """The coordinates (and input weightings) of many grid pairs in one
    device buffer, with the descriptor array rg_build_conservative_2d_batched
    takes.  Keep it alive while builds read it."""

    def __init__(self, pairs):
        arrays, shapes = [], []
        for item in pairs:
            gi, go = item[0], item[1]
            wi = item[2] if len(item) > 2 else None
            xi = np.ascontiguousarray(gi[0], dtype=np.float64)
            yi = np.ascontiguousarray(gi[1], dtype=np.float64)
            xo = np.ascontiguousarray(go[0], dtype=np.float64)
            yo = np.ascontiguousarray(go[1], dtype=np.float64)
            if xi.ndim != 2 or xo.ndim != 2 or xi.shape != yi.shape or xo.shape != yo.shape:
                raise ValueError("2-D grids need two coordinate arrays of one 2-D shape")
            if wi is not None:
                wi = np.ascontiguousarray(wi, dtype=np.float64)
                if wi.shape != (xi.shape[0] - 1, xi.shape[1] - 1):
                    raise ValueError(
                        f"weights_input has shape {wi.shape}, expected the input cell "
                        f"shape {(xi.shape[0] - 1, xi.shape[1] - 1)}"
                    )
            arrays.append((xi, yi, xo, yo, wi))
            shapes.append((xi.shape, xo.shape))
        sizes = [sum(a.size for a in arr if a is not None) for arr in arrays]
        flat = np.empty(max(sum(sizes), 1), dtype=np.float64)
        offsets = []
        pos = 0
        for arr in arrays:
            offs = []
            for a in arr:
                if a is None:
                    offs.append(None)
                    continue
                flat[pos : pos + a.size] = a.ravel()
                offs.append(pos)
                pos += a.size
            offsets.append(offs)
        self.buffer = DeviceArray.from_host(flat)
        self._layout = (shapes, offsets, flat.size)
        self.n = len(arrays)
        self.desc = (rg_pair_desc * max(self.n, 1))()
        base = self.buffer.ptr
        for k, ((si, so), offs) in enumerate(zip(shapes, offsets)):
            d = self.desc[k]
            d.xi, d.yi = base + 8 * offs[0], base + 8 * offs[1]
            d.nxi, d.nyi = si
            d.xo, d.yo = base + 8 * offs[2], base + 8 * offs[3]
            d.nxo, d.nyo = so
            d.wi = None if offs[4] is None else base + 8 * offs[4]


    def update(self, pairs) -> None:
        """Overwrite the coordinates (and weightings) in place with pairs of
        the same shapes: builds reading this buffer see the new values."""
        shapes, offsets, size = self._layout
        flat = np.empty(max(size, 1), dtype=np.float64)
        if len(pairs) != len(shapes):
            raise ValueError("a different number of pairs")
        for item, (si, so), offs in zip(pairs, shapes, offsets):
            gi, go = item[0], item[1]
            wi = item[2] if len(item) > 2 else None
            arrs = (gi[0], gi[1], go[0], go[1], wi)
            for a, o, sh in zip(arrs, offs, (si, si, so, so, None)):
                if o is None:
                    if a is not None:
                        raise ValueError("a weighting where the packed pair has none")
                    continue
                a = np.ascontiguousarray(a, dtype=np.float64)
                if sh is not None and a.shape != tuple(sh):
                    raise ValueError(f"shape {a.shape} differs from the packed {tuple(sh)}")
                flat[o: o + a.size] = a.ravel()
        self.buffer.copy_from_host(flat[: max(size, 1)] if size else flat)


def pack_pairs(pairs) -> PackedPairs:
    """Copy ``[(grid_input, grid_output[, weights_input]), ...]`` (host
    arrays, already normalised and perturbed) to the device in one transfer."""
    return PackedPairs(pairs)


def build_2d_batched(pairs, edges_per_lane: int = 0, block: int = 0,
                     timings: bool = False, stage_per_edge: int = 0,
                     stream: int = 0, flags: int = 0) -> list[Weights]:
    """Build many independent grid pairs together (one launch per phase for
    the whole batch).  ``pairs`` is a PackedPairs or a sequence accepted by
    pack_pairs.  Returns one handle per pair, each bit-identical to the
    pair's own build_2d.  ``flags`` adds ``RG_TUNE_*`` bits (test hooks)."""
    packed = pairs if isinstance(pairs, PackedPairs) else PackedPairs(pairs)
    if packed.n == 0:
        return []
    t = rg_tuning(int(edges_per_lane), int(block),
                  (RG_TUNE_TIMINGS if timings else 0) | int(flags), int(stage_per_edge))
    hs = (ctypes.c_void_p * packed.n)()
    check(lib().rg_build_conservative_2d_batched(packed.desc, packed.n, ctypes.byref(t), hs,
                                                 stream))
    return [Weights(h) for h in hs]
