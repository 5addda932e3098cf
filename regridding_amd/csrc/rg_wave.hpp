// rg_wave.hpp -- wavefront-cooperative helpers (64 lanes on gfx950).
#pragma once

#include "rg_sweep.hpp"

namespace rg {

// The calling wavefront's index in its block, and in the grid (blocks of
// whole wavefronts), as wave-uniform values.  The compiler cannot tell that
// threadIdx.x >> 6 is the same on every lane; whatever a kernel derives from
// it would sit in vector registers and be read with vector loads, one
// dependent round trip after another.  Read this way it sits in a scalar
// register, and the per-wavefront tables it indexes are read with scalar
// loads.
__device__ __forceinline__ int block_wave()
{
    return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
}

__device__ __forceinline__ int64_t grid_wave()
{
    return (int64_t)blockIdx.x * (blockDim.x >> 6) + block_wave();
}

__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ inline int64_t wave_min(int64_t v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        int64_t w = __shfl_xor(v, o, 64);
        v = w < v ? w : v;
    }
    return v;
}

// Lowest flat index of a cell containing (px, py), cooperatively over the
// 64 lanes of a wavefront (index_of_point_brute, _grids.py:223-279, returns
// the first cell in row-major order); -1 if none.
__device__ inline int64_t wave_brute(const double* __restrict__ x,
                                     const double* __restrict__ y, int64_t nx,
                                     int64_t ny, double px, double py, int lane)
{
    const int64_t ncy = ny - 1, ncells = (nx - 1) * ncy;
    int64_t best = INT64_MAX;
    for (int64_t c = lane; c < ncells; c += 64)
        if (cell_contains(x, y, ny, c / ncy, c % ncy, px, py)) {
            best = c;
            break;
        }
    best = wave_min(best);
    return best == INT64_MAX ? -1 : best;
}

}  // namespace rg
