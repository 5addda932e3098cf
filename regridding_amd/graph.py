"""
One hipGraph for a batched build and apply (``rg_graph_*``).

The batched build of many grid pairs (one launch per phase) followed by the
FP64 apply of each pair's weights to its values, captured once and then
replayed with a single ``hipGraphLaunch``.  The coordinates and values are
read at the device addresses given at creation: ``PackedPairs.update`` and
``set_values`` change them in place (same shapes) between replays.
"""

from __future__ import annotations

import ctypes

import numpy as np

from ._conservative import PackedPairs
from ._lib import DeviceArray, RGError, Weights, check, lib, rg_graph_job, rg_tuning

__all__ = ["BuildApplyGraph"]

FLAG_NAMES = {1: "triples", 2: "overflow triples", 4: "weights", 8: "sweep failed",
              16: "emission mismatch"}


class BuildApplyGraph:
    """
    ``pairs``: a ``PackedPairs`` (kept alive; its buffer is what replays
    read).  ``values``: per pair a (batch, n_in) float64 host array (batch
    may be 0; None: build only), copied to the device once; or
    ``device_jobs``: per pair (input pointer, output pointer, batch) of
    device arrays the caller owns.  ``slack``: extra room over the measured
    triple and weight counts, for replays whose coordinates produce more of
    them.  ``edges_per_lane``, ``block``, ``stage_per_edge``: the build's
    launch shape and staging (0: defaults).
    """

    def __init__(self, pairs: PackedPairs, values=None, slack: float = 0.25,
                 edges_per_lane: int = 0, block: int = 0, stream: int = 0,
                 device_jobs=None, stage_per_edge: int = 0):
        self.pairs = pairs
        L = lib()
        n = pairs.n
        self.n_in, self.n_out, self.batch = [], [], []
        for k in range(n):
            d = pairs.desc[k]
            self.n_in.append((d.nxi - 1) * (d.nyi - 1))
            self.n_out.append((d.nxo - 1) * (d.nyo - 1))
        t = rg_tuning(int(edges_per_lane), int(block), 0, int(stage_per_edge))
        jobs = (rg_graph_job * n)()
        self.stream = stream
        if device_jobs is not None:
            # (input pointer, output pointer, batch) per pair: device arrays
            # the caller owns and keeps alive while the graph exists
            self.x = self.y = None
            if len(device_jobs) != n:
                raise ValueError(f"{len(device_jobs)} jobs for {n} pairs")
            for k, (xp, yp, b) in enumerate(device_jobs):
                jobs[k].values_input, jobs[k].values_output, jobs[k].batch = xp, yp, int(b)
                self.batch.append(int(b))
            self._create(pairs, t, jobs, slack, stream)
            return
        if values is None:  # build only
            values = [None] * n
        if len(values) != n:
            raise ValueError(f"{len(values)} value arrays for {n} pairs")
        x_parts, self._x_off, self._y_off = [], [], []
        xo = yo = 0
        for k, v in enumerate(values):
            v = np.zeros((0, self.n_in[k])) if v is None else np.asarray(v, dtype=np.float64)
            v = v.reshape(v.shape[0], -1) if v.size else v.reshape(-1, self.n_in[k])
            if v.shape[1] != self.n_in[k]:
                raise ValueError(f"values of pair {k} have {v.shape[1]} cells, not {self.n_in[k]}")
            self.batch.append(v.shape[0])
            x_parts.append(v.ravel())
            self._x_off.append(xo)
            self._y_off.append(yo)
            xo += v.size
            yo += v.shape[0] * self.n_out[k]
        self.x = DeviceArray.from_host(np.concatenate(x_parts) if xo else np.zeros(1))
        self.y = DeviceArray.empty((max(yo, 1),), np.float64)
        for k in range(n):
            jobs[k].values_input = self.x.ptr + 8 * self._x_off[k]
            jobs[k].values_output = self.y.ptr + 8 * self._y_off[k]
            jobs[k].batch = self.batch[k]
        self._create(pairs, t, jobs, slack, stream)

    def _create(self, pairs, t, jobs, slack, stream):
        h = ctypes.c_void_p()
        self.ptr = None
        check(lib().rg_graph_create(pairs.desc, pairs.n, ctypes.byref(t), jobs, float(slack),
                                    ctypes.byref(h), stream))
        self.ptr = h.value

    def set_values(self, values) -> None:
        """New values of the same shapes for the next replays."""
        if self.x is None:
            raise ValueError("the graph applies to the caller's device arrays")
        parts = [np.asarray(v, dtype=np.float64).ravel() for v in values]
        flat = np.concatenate(parts) if parts else np.zeros(1)
        self.x.copy_from_host(flat if flat.size else np.zeros(1))

    def launch(self, stream: int | None = None) -> None:
        check(lib().rg_graph_launch(self.ptr, self.stream if stream is None else stream))

    def status(self) -> int:
        """Waits for the stream; 0 when the last replay fit its capacities."""
        f = ctypes.c_int32()
        check(lib().rg_graph_status(self.ptr, ctypes.byref(f), self.stream))
        return f.value

    def check(self) -> None:
        f = self.status()
        if f:
            names = ", ".join(v for b, v in FLAG_NAMES.items() if f & b)
            raise RGError(6, f"the graph's replay exceeded its capacity ({names}): "
                             "create it again with more slack")

    def outputs(self) -> list[np.ndarray]:
        """Per pair the (batch, n_out) results of the last replay."""
        if self.y is None:
            raise ValueError("the graph applies to the caller's device arrays")
        y = self.y.to_host(self.stream)
        return [y[o: o + b * m].reshape(b, m)
                for o, b, m in zip(self._y_off, self.batch, self.n_out)]

    def weights(self, k: int) -> Weights:
        """A copy of pair k's weights from the last replay."""
        h = ctypes.c_void_p()
        check(lib().rg_graph_weights(self.ptr, int(k), ctypes.byref(h), self.stream))
        return Weights(h.value)

    def capacity(self) -> dict:
        vals = [ctypes.c_int64() for _ in range(4)]
        check(lib().rg_graph_capacity(self.ptr, *[ctypes.byref(v) for v in vals]))
        return dict(zip(("triples", "overflow", "weights", "workspace_bytes"),
                        (v.value for v in vals)))

    def close(self) -> None:
        if getattr(self, "ptr", None):
            check(lib().rg_graph_destroy(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
