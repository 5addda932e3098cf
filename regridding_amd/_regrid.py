"""
``regrid_from_weights()`` and ``regrid()``: the drop-in apply
(``regridding/_regrid/_regrid_from_weights.py:12-162``,
``regridding/_regrid/_regrid.py:12-149``).

The host part (axis normalisation, broadcasting, moving the resampled axes
last and flattening the orthogonal ones) follows the reference step by step;
the loop ``out[d][i_out] += w * in[d][i_in]`` runs on the GPU as a CSR
gather (``rg_apply_f64``) that accumulates every output in the reference's
order, so the result is bit-identical.  Orthogonal elements that share one
weight matrix (the usual broadcast case) are applied in one batched launch.
"""

from __future__ import annotations

import ctypes
from typing import Literal, Sequence

import numpy as np

from . import _cache, _util
from ._lib import DeviceArray, Weights, check, lib

__all__ = ["regrid_from_weights", "regrid", "import_weights", "prepare_apply", "finish_apply"]


def import_weights(indices_input, indices_output, values, n_in: int, n_out: int) -> Weights:
    """Upload arbitrary triples (any order) as a device handle.  Negative
    indices wrap around, as the reference's numba indexing does (the 1-D
    builders emit them for descending grids)."""
    ii = np.ascontiguousarray(indices_input, dtype=np.int64)
    io = np.ascontiguousarray(indices_output, dtype=np.int64)
    if ii.size and ii.min() < 0:
        ii = np.where(ii < 0, ii + int(n_in), ii)
    if io.size and io.min() < 0:
        io = np.where(io < 0, io + int(n_out), io)
    w = np.ascontiguousarray(np.asarray(values), dtype=np.float64)
    if not (ii.shape == io.shape == w.shape and ii.ndim == 1):
        raise ValueError("weight arrays must be 1-D and of one length")
    di, do, dw = (DeviceArray.from_host(a) for a in (ii, io, w))
    h = ctypes.c_void_p()
    check(lib().rg_weights_import(di.ptr, do.ptr, dw.ptr, ii.size, int(n_in), int(n_out),
                                  ctypes.byref(h), 0))
    return Weights(h.value)


def prepare_apply(
    weights,
    shape_input,
    shape_output,
    values_input,
    values_output=None,
    axis_input=None,
    axis_output=None,
):
    """
    The host half of ``regrid_from_weights`` before the compiled loop
    (``_regrid_from_weights.py:62-144``).

    Returns a dict with ``flat`` (the per-element weight triples, length D),
    ``values_input`` (D, n_in) float64 C-contiguous, ``values_output`` the
    (D, *cells_out) zero-filled buffer the result goes to, plus the pieces
    :func:`finish_apply` needs.
    """
    unit = getattr(values_input, "unit", None)

    ndim_input = len(shape_input)
    ndim_output = len(shape_output)

    axis_input = _util.normalize_axis(axis_input, ndim=ndim_input)
    axis_output = _util.normalize_axis(axis_output, ndim=ndim_output)

    orth_in = tuple(
        shape_input[i] for i in _util.normalize_axis(None, ndim_input) if i not in axis_input
    )
    orth_out = tuple(
        shape_output[i] for i in _util.normalize_axis(None, ndim_output) if i not in axis_output
    )
    if np.ndim(values_input) > 0:
        orth_values = tuple(
            values_input.shape[i]
            for i in _util.normalize_axis(None, values_input.ndim)
            if i not in axis_input
        )
    else:
        orth_values = ()
    shape_orthogonal = np.broadcast_shapes(orth_in, orth_out, orth_values)

    axis_input = tuple(sorted(axis_input))
    axis_output = tuple(sorted(axis_output))

    full_in = list(reversed(shape_orthogonal))
    for ax in reversed(axis_input):
        full_in.insert(~ax, shape_input[ax])
    full_in = tuple(reversed(full_in))

    full_out = list(reversed(shape_orthogonal))
    for ax in reversed(axis_output):
        full_out.insert(~ax, shape_output[ax])
    full_out = tuple(reversed(full_out))

    weights = np.broadcast_to(np.array(weights), shape_orthogonal, subok=True)
    values_input = np.broadcast_to(values_input, full_in, subok=True)

    if values_output is None:
        values_output = np.zeros_like(values_input, shape=full_out, dtype=float)
    else:
        if values_output.shape != full_out:
            raise ValueError(f"{values_output.shape=} should be equal to {full_out}")
        values_output.fill(0)

    axis_input_last = ~np.arange(len(axis_input))[::-1]
    axis_output_last = ~np.arange(len(axis_output))[::-1]
    cells_in = tuple(full_in[ax] for ax in axis_input)
    cells_out = tuple(full_out[ax] for ax in axis_output)

    values_input = np.moveaxis(values_input, axis_input, axis_input_last)
    values_output = np.moveaxis(values_output, axis_output, axis_output_last)
    shape_output_moved = values_output.shape

    values_input = values_input.reshape(-1, *cells_in)
    values_output = values_output.reshape(-1, *cells_out)

    flat = weights.reshape(-1)
    unit_weights = getattr(flat[0][2], "unit", None) if flat.size else None

    values_input = np.ascontiguousarray(np.asarray(values_input), dtype=np.float64)
    values_output = np.ascontiguousarray(values_output)

    return dict(
        flat=flat,
        values_input=values_input.reshape(values_input.shape[0], -1),
        values_output=values_output,
        n_in=int(np.prod(cells_in)),
        n_out=int(np.prod(cells_out)),
        shape_output_moved=shape_output_moved,
        axis_output=axis_output,
        axis_output_last=axis_output_last,
        unit=unit,
        unit_weights=unit_weights,
    )


def finish_apply(prep: dict, result: np.ndarray) -> np.ndarray:
    """Write `result` (D, n_out) into the output buffer and restore the
    caller's axis order and units (``_regrid_from_weights.py:152-162``)."""
    out = prep["values_output"]
    out.reshape(out.shape[0], -1)[...] = result
    out = out.reshape(*prep["shape_output_moved"])
    out = np.moveaxis(out, prep["axis_output_last"], prep["axis_output"])
    unit, unit_weights = prep["unit"], prep["unit_weights"]
    if unit_weights is not None:
        unit = unit_weights if unit is None else unit * unit_weights
    if unit is None:
        return out
    return out << unit


def _handle_for(element, n_in: int, n_out: int) -> tuple[Weights, bool]:
    ii, io, w = element
    w = np.asarray(w)
    if isinstance(ii, np.ndarray) and isinstance(io, np.ndarray):
        hit = _cache.lookup(ii, io, w)
        if hit is not None:
            return hit
    return import_weights(ii, io, w, n_in, n_out), False


def _block_diagonal(flat, n_in: int, n_out: int):
    """``(handle, transposed)`` when element d of ``flat`` is block d of one
    block-diagonal handle for every d (the 1-D builders' output applied as
    returned), else None."""
    D = flat.size
    if D == 0:
        return None
    first = None
    for d in range(D):
        ii, io, w = flat[d]
        if not (isinstance(ii, np.ndarray) and isinstance(io, np.ndarray)):
            return None
        hit = _cache.lookup_part(ii, io, np.asarray(w))
        if hit is None or hit[2] is None:
            return None
        handle, transposed, (k, nparts, bi, bo) = hit
        if transposed:
            bi, bo = bo, bi
        if k != d or nparts != D or (bi, bo) != (n_in, n_out):
            return None
        if first is None:
            first = (handle, transposed)
        elif handle is not first[0] or transposed != first[1]:
            return None
    return first


def apply_device(handle: Weights, transposed: bool, x: np.ndarray, n_out: int) -> np.ndarray:
    """(D, n_in) host float64 -> (D, n_out) host float64 through the GPU."""
    D = x.shape[0]
    out = np.empty((D, n_out))
    if D == 0 or n_out == 0:
        return out
    dx = DeviceArray.from_host(x)
    dy = DeviceArray.empty((D, n_out), np.float64)
    check(lib().rg_apply_f64(handle.ptr, int(transposed), dx.ptr, dy.ptr, D, 0))
    return dy.to_host(out=out)


def _check_sizes(handle: Weights, transposed: bool, n_in: int, n_out: int) -> None:
    hin, hout = (handle.n_out, handle.n_in) if transposed else (handle.n_in, handle.n_out)
    if (hin, hout) != (n_in, n_out):
        raise ValueError(
            f"weights relate {hin} input cells to {hout} output cells, but the "
            f"values have {n_in} input cells and {n_out} output cells"
        )


def regrid_from_weights(
    weights: np.ndarray,
    shape_input: tuple[int, ...],
    shape_output: tuple[int, ...],
    values_input: np.ndarray,
    values_output: None | np.ndarray = None,
    axis_input: None | int | Sequence[int] = None,
    axis_output: None | int | Sequence[int] = None,
) -> np.ndarray:
    """
    Apply weights from :func:`weights` to an array of values.

    Same signature, broadcasting and result (float64, bit-identical) as
    ``regridding.regrid_from_weights``.
    """
    prep = prepare_apply(weights, shape_input, shape_output, values_input,
                         values_output, axis_input, axis_output)
    flat = prep["flat"]
    x = prep["values_input"]
    n_in, n_out = prep["n_in"], prep["n_out"]
    D = x.shape[0]
    whole = _block_diagonal(flat, n_in, n_out)
    if whole is not None:
        # every element is its own block of one handle, in order: one launch
        handle, transposed = whole
        y = apply_device(handle, transposed, x.reshape(1, D * n_in), D * n_out)
        return finish_apply(prep, y.reshape(D, n_out))
    result = np.zeros((D, n_out))
    # group the elements by weight object: broadcast weights are one object
    groups: dict[int, list[int]] = {}
    for d in range(D):
        groups.setdefault(id(flat[d]), []).append(d)
    for ds in groups.values():
        handle, transposed = _handle_for(flat[ds[0]], n_in, n_out)
        _check_sizes(handle, transposed, n_in, n_out)
        idx = np.asarray(ds)
        block = x if len(ds) == D else np.ascontiguousarray(x[idx])
        y = apply_device(handle, transposed, block, n_out)
        if len(ds) == D:
            result = y
        else:
            result[idx] = y
    return finish_apply(prep, result)


def regrid(
    coordinates_input: tuple[np.ndarray, ...],
    coordinates_output: tuple[np.ndarray, ...],
    values_input: np.ndarray,
    values_output: None | np.ndarray = None,
    axis_input: None | int | Sequence[int] = None,
    axis_output: None | int | Sequence[int] = None,
    method: Literal["multilinear", "conservative"] = "multilinear",
    bounds: Literal["extrapolate", "nan", "raise"] = "extrapolate",
    perturb: None | bool = None,
    seed: "None | int | np.random.Generator" = _util.SEED_DEFAULT,
) -> np.ndarray:
    """
    Regrid values from one grid to another in one call: :func:`weights`
    then :func:`regrid_from_weights` (``regridding/_regrid/_regrid.py:130-148``).
    The weights stay on the device in between.
    """
    from ._weights import weights as _weights

    w = _weights(
        coordinates_input=coordinates_input,
        coordinates_output=coordinates_output,
        axis_input=axis_input,
        axis_output=axis_output,
        method=method,
        bounds=bounds,
        perturb=perturb,
        seed=seed,
    )
    return regrid_from_weights(
        weights=w[0],
        shape_input=w[1],
        shape_output=w[2],
        values_input=values_input,
        values_output=values_output,
        axis_input=axis_input,
        axis_output=axis_output,
    )
