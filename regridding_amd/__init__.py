"""
regridding_amd -- conservative regridding between logically-rectangular
curvilinear grids on AMD Instinct MI355X (gfx950).

A drop-in for the hot path of ``regridding`` (sun-data/regridding): the same
entry points, signatures, broadcasting rules, errors and results, with the
2-D conservative weight build, the cell search and the weight application
running as hand-written HIP kernels behind a C ABI (``librg_hip.so``,
``include/regridding_amd.h``).
"""

from ._util import SEED_DEFAULT
from ._weights import weights
from ._regrid import regrid_from_weights, regrid
from ._transpose import transpose_weights, transpose_weights_conservative
from ._find_indices import find_indices
from ._fill import fill
from ._interp import ndarray_linear_interpolation
from . import weights_io

__all__ = [
    "SEED_DEFAULT",
    "weights",
    "regrid_from_weights",
    "regrid",
    "transpose_weights",
    "transpose_weights_conservative",
    "find_indices",
    "fill",
    "ndarray_linear_interpolation",
]
