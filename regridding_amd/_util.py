"""
Host-side argument normalisation shared by the drop-in entry points.

This is the part of the conservative build that stays on the host: axis
normalisation, broadcasting of the coordinate arrays and the seeded Gaussian
perturbation of the output grid.  The perturbation must be drawn exactly as
the reference draws it (``regridding/_util.py:121-129``), because every later
discrete decision of the sweep depends on the perturbed coordinates.
"""

from __future__ import annotations

from typing import Sequence

import numpy as np

__all__ = [
    "SEED_DEFAULT",
    "normalize_axis",
    "normalize_coordinates",
]

SEED_DEFAULT = 42
"""Default seed of the output-grid perturbation (``regridding/_util.py:4``)."""

PERTURB_EPSILON = 1e-9
"""Relative scale of the perturbation (``regridding/_util.py:122``)."""


def normalize_axis(
    axis: None | int | Sequence[int],
    ndim: int,
) -> tuple[int, ...]:
    """
    Return `axis` as a tuple of negative axis numbers.

    ``None`` selects every axis.  Follows ``regridding/_util.py:12-20``:
    out-of-range or repeated axes raise numpy's ``AxisError`` /
    ``ValueError``.
    """
    if axis is None:
        axis = tuple(range(ndim))
    positive = np.lib.array_utils.normalize_axis_tuple(axis, ndim=ndim)
    return tuple(int(a) - ndim for a in positive)


def _shape_with_axes(
    shape_orthogonal: tuple[int, ...],
    axes: tuple[int, ...],
    shape_source: tuple[int, ...],
) -> tuple[int, ...]:
    """
    Insert the sizes ``shape_source[ax]`` for each (negative) axis of `axes`
    into `shape_orthogonal`, counting positions from the right.

    Equivalent to the reversed-list insertion in
    ``regridding/_util.py:104-112``; `axes` must be sorted descending.
    """
    reversed_shape = list(reversed(shape_orthogonal))
    for ax in axes:
        reversed_shape.insert(~ax, shape_source[ax])
    return tuple(reversed(reversed_shape))


def normalize_coordinates(
    coordinates_input: np.ndarray | tuple[np.ndarray, ...],
    coordinates_output: np.ndarray | tuple[np.ndarray, ...],
    axis_input: None | int | Sequence[int] = None,
    axis_output: None | int | Sequence[int] = None,
    perturb: bool = False,
    seed: "None | int | np.random.Generator" = SEED_DEFAULT,
) -> tuple[
    tuple[np.ndarray, ...],
    tuple[np.ndarray, ...],
    tuple[int, ...],
    tuple[int, ...],
    tuple[int, ...],
    tuple[int, ...],
    tuple[int, ...],
]:
    """
    Broadcast, validate and (optionally) perturb the coordinate arrays.

    Returns ``(coordinates_input, coordinates_output, axis_input,
    axis_output, shape_coordinates_input, shape_coordinates_output,
    shape_orthogonal)`` with the same meaning, order and error behaviour as
    ``regridding/_util.py:23-138``:

    * the axes come back negative and sorted descending;
    * both coordinate tuples are broadcast to the full vertex shapes
      (orthogonal axes broadcast between input and output);
    * with `perturb`, each output coordinate ``c`` (x first, then y) is
      replaced by ``rng.normal(c, ptp(c over the resampled axes) * 1e-9)``
      with ``rng = numpy.random.default_rng(seed)``.
    """
    if isinstance(coordinates_input, np.ndarray):
        coordinates_input = (coordinates_input,)
    if isinstance(coordinates_output, np.ndarray):
        coordinates_output = (coordinates_output,)

    # united coordinates are recognised by duck typing: the input is
    # converted to the output's unit before any comparison
    converted = []
    for c_in, c_out in zip(coordinates_input, coordinates_output):
        unit = getattr(c_out, "unit", None)
        converted.append(c_in << unit if unit is not None else c_in)
    coordinates_input = tuple(converted)

    shape_in = np.broadcast(*coordinates_input).shape
    shape_out = np.broadcast(*coordinates_output).shape

    axis_input = tuple(sorted(normalize_axis(axis_input, len(shape_in)), reverse=True))
    axis_output = tuple(
        sorted(normalize_axis(axis_output, len(shape_out)), reverse=True)
    )

    if len(axis_output) != len(axis_input):
        raise ValueError(
            f"The number of axes in `axis_output`, {axis_output}, "
            f"must match the number of axes in `axis_input`, {axis_input}"
        )
    if len(coordinates_input) != len(axis_input):
        raise ValueError(
            f"The number of elements in `coordinates_input`, "
            f"{len(coordinates_input)}, should match the number of axes in "
            f"`axis_input`, {axis_input}"
        )
    if len(coordinates_output) != len(coordinates_input):
        raise ValueError(
            f"The number of elements in `coordinates_output`, "
            f"{len(coordinates_output)}, should match the number of elements "
            f"in `coordinates_input`, {len(coordinates_input)}"
        )

    orth_in = tuple(
        shape_in[ax] for ax in normalize_axis(None, len(shape_in)) if ax not in axis_input
    )
    orth_out = tuple(
        shape_out[ax]
        for ax in normalize_axis(None, len(shape_out))
        if ax not in axis_output
    )
    shape_orthogonal = np.broadcast_shapes(orth_in, orth_out)

    full_in = _shape_with_axes(shape_orthogonal, axis_input, shape_in)
    full_out = _shape_with_axes(shape_orthogonal, axis_output, shape_out)

    coordinates_input = tuple(np.broadcast_to(c, full_in) for c in coordinates_input)
    coordinates_output = tuple(
        np.broadcast_to(c, full_out) for c in coordinates_output
    )

    if perturb:
        rng = np.random.default_rng(seed)
        jittered = []
        for c in coordinates_output:
            scale = np.ptp(c, axis=axis_output, keepdims=True) * PERTURB_EPSILON
            jittered.append(rng.normal(c, scale))
        coordinates_output = tuple(jittered)

    return (
        coordinates_input,
        coordinates_output,
        axis_input,
        axis_output,
        shape_in,
        shape_out,
        shape_orthogonal,
    )


def resolve_perturb(
    perturb: None | bool,
    coordinates_input: np.ndarray | tuple[np.ndarray, ...],
) -> bool:
    """
    ``perturb=None`` means "perturb when there are two or more coordinates"
    (``regridding/_weights/_weights_conservative.py:21-25``).
    """
    if perturb is not None:
        return bool(perturb)
    if isinstance(coordinates_input, np.ndarray):
        return False
    return len(coordinates_input) > 1


def orthogonal_slice(
    index: tuple[int, ...],
    axes: tuple[int, ...],
) -> tuple:
    """
    Build the index that selects orthogonal element `index` of a full
    vertex (or cell) array and keeps the resampled `axes` (negative, sorted
    descending) whole, as ``_weights_conservative.py:110-120`` does.
    """
    reversed_index: list = list(reversed(index))
    for ax in axes:
        reversed_index.insert(~ax, slice(None))
    return tuple(reversed(reversed_index))
