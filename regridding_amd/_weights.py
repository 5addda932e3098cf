"""
``weights()``: the drop-in dispatcher (``regridding/_weights/_weights.py:13-196``).
"""

from __future__ import annotations

from typing import Literal, Sequence

import numpy as np

from . import _cache, _line, _util
from ._conservative import build_2d_batched

__all__ = ["weights", "conservative_weights"]


def conservative_weights(
    coordinates_input,
    coordinates_output,
    axis_input=None,
    axis_output=None,
    weights_input=None,
    perturb=None,
    seed=_util.SEED_DEFAULT,
):
    """
    ``_weights_conservative`` (``_weights_conservative.py:11-146``) followed
    by ``_weights_to_arrays`` (``_weights_arrays.py:8-41``): normalise, then
    build every orthogonal element on the device -- 2-D elements in one
    batched build, 1-D elements in one block-diagonal sweep -- and return
    the coalesced per-element triples (registered with their device handle
    for :func:`regrid_from_weights`).
    """
    perturb = _util.resolve_perturb(perturb, coordinates_input)
    (
        coordinates_input,
        coordinates_output,
        axis_input,
        axis_output,
        shape_input,
        shape_output,
        shape_orthogonal,
    ) = _util.normalize_coordinates(
        coordinates_input=coordinates_input,
        coordinates_output=coordinates_output,
        axis_input=axis_input,
        axis_output=axis_output,
        perturb=perturb,
        seed=seed,
    )

    shape_values_input = list(shape_input)
    for ax in axis_input:
        shape_values_input[ax] -= 1
    shape_values_input = tuple(shape_values_input)

    shape_values_output = list(shape_output)
    for ax in axis_output:
        shape_values_output[ax] -= 1
    shape_values_output = tuple(shape_values_output)

    if weights_input is not None:
        weights_input = np.broadcast_to(weights_input, shape_values_input)

    if len(axis_input) == 1:
        (x_input,) = coordinates_input
        (x_output,) = coordinates_output
        x_input = np.moveaxis(x_input, axis_input, ~0)
        x_output = np.moveaxis(x_output, axis_output, ~0)
        x_input = x_input.reshape(-1, x_input.shape[~0])
        x_output = x_output.reshape(-1, x_output.shape[~0])
        if weights_input is not None:
            weights_input = np.moveaxis(weights_input, axis_input, ~0)
            weights_input = weights_input.reshape(-1, weights_input.shape[~0])
        result = _line.conservative_1d(x_input, x_output, weights_input, shape_orthogonal)
        return result, shape_values_input, shape_values_output
    if len(axis_input) != 2:
        raise NotImplementedError(
            "Regridding operations greater than 2D are not supported"
        )

    x_in, y_in = coordinates_input
    x_out, y_out = coordinates_output
    items = []
    for index in np.ndindex(*shape_orthogonal):
        sl_in = _util.orthogonal_slice(index, axis_input)
        sl_out = _util.orthogonal_slice(index, axis_output)
        # the reference indexes weights_input with the bare orthogonal index
        # (_weights_conservative.py:126)
        wi = weights_input[index] if weights_input is not None else None
        items.append(((x_in[sl_in], y_in[sl_in]), (x_out[sl_out], y_out[sl_out]), wi))
    handles = build_2d_batched(items)
    result = np.empty(len(handles), dtype=object)
    for k, h in enumerate(handles):
        ii, io, w = h.export()
        _cache.register(ii, io, w, h)
        result[k] = (ii, io, w)
    return result.reshape(shape_orthogonal), shape_values_input, shape_values_output


def weights(
    coordinates_input: tuple[np.ndarray, ...],
    coordinates_output: tuple[np.ndarray, ...],
    axis_input: None | int | Sequence[int] = None,
    axis_output: None | int | Sequence[int] = None,
    weights_input: None | np.ndarray = None,
    method: Literal["multilinear", "conservative"] = "multilinear",
    bounds: Literal["extrapolate", "nan", "raise"] = "extrapolate",
    perturb: None | bool = None,
    seed: "None | int | np.random.Generator" = _util.SEED_DEFAULT,
) -> tuple[np.ndarray, tuple[int, ...], tuple[int, ...]]:
    """
    Build the sparse weights relating two logically-rectangular grids.

    Same signature, return value and errors as ``regridding.weights``
    (``regridding/_weights/_weights.py:13-196``): an object array over the
    orthogonal axes whose elements are ``(indices_input, indices_output,
    values)`` flat arrays sorted by ``(input, output)`` with unique pairs,
    plus the input and output value shapes.  The builds run on the GPU; the
    weights stay resident there for :func:`regrid_from_weights`, which uses
    the resident copy while the returned (writable) arrays still hold the
    bytes they were returned with, and imports them again once edited.
    """
    unit_weights = getattr(weights_input, "unit", None)
    if unit_weights is not None:
        weights_input = getattr(weights_input, "value")

    if method == "multilinear":
        result, shape_values_input, shape_values_output = _line.multilinear_weights(
            coordinates_input=coordinates_input,
            coordinates_output=coordinates_output,
            axis_input=axis_input,
            axis_output=axis_output,
            weights_input=weights_input,
            bounds=bounds,
            perturb=perturb,
            seed=seed,
        )
    elif method == "conservative":
        result, shape_values_input, shape_values_output = conservative_weights(
            coordinates_input=coordinates_input,
            coordinates_output=coordinates_output,
            axis_input=axis_input,
            axis_output=axis_output,
            weights_input=weights_input,
            perturb=perturb,
            seed=seed,
        )
    else:
        raise ValueError(f"unrecognized method '{method}'")

    if unit_weights is not None:
        flat = result.reshape(-1)
        for k in range(flat.size):
            ii, io, w = flat[k]
            flat[k] = (ii, io, w << unit_weights)
    return result, shape_values_input, shape_values_output
