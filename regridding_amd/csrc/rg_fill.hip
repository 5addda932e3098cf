// rg_fill.hip -- the reference's companions of the regrid: filling missing
// values by red-black Gauss-Seidel relaxation and linear interpolation of an
// array at fractional indices.
//
// Red-black Gauss-Seidel (regridding/_fill/_gauss_seidel.py:83-139).  The
// reference sweeps each colour row-major with periodic neighbours.  Within
// one colour a cell's neighbours have the other colour -- except across the
// periodic wrap when a dimension is odd: cells (j, 0) and (j, nx-1) then
// share a colour (likewise (0, i) and (ny-1, i)), and the sequential sweep
// updates the later one from the already-updated earlier one.  Every such
// pair has exactly one "late" member (last column for odd nx, last row for
// odd ny), so a colour is applied in steps that give the sequential sweep's
// result exactly:
//   1. every non-late cell of the colour, in parallel: a same-colour
//      neighbour of one is late, and still holds its old value in the
//      sequential sweep too;
//   2. every late cell but the corner (ny-1, nx-1), in parallel: their
//      same-colour neighbours are non-late (already new, as in the sweep)
//      or the corner (still old, as in the sweep);
//   3. the corner, when both dimensions are odd: its same-colour neighbours
//      (0, nx-1) and (ny-1, 0) are late and already new.
// A slice that fits in LDS runs every iteration in one workgroup without
// leaving LDS (k_gs_lds); larger slices take one launch per step.
//
// Linear interpolation (regridding/_interp_ndarray.py:172-269): one lane per
// output point, clamped linear / bilinear weights.
#include <hip/hip_runtime.h>

#include "rg_internal.hpp"

namespace rg {

namespace {

struct GsCoef {
    double dxxinv, dyyinv, dcent;
};

// _gauss_seidel.py:110-122
GsCoef gs_coef(int64_t nx, int64_t ny)
{
    const double xmin = -1, xmax = 1, ymin = -1, ymax = 1;
    const double dx = (xmax - xmin) / (double)(nx - 1);
    const double dy = (ymax - ymin) / (double)(ny - 1);
    GsCoef c;
    c.dxxinv = 1 / (dx * dx);
    c.dyyinv = 1 / (dy * dy);
    c.dcent = 1 / (2 * (c.dxxinv + c.dyyinv));
    return c;
}

// _gauss_seidel.py:124-139: one cell
__device__ inline void gs_cell(double* __restrict__ a, int64_t nx, int64_t ny,
                               int64_t j, int64_t i, GsCoef c)
{
    const int64_t i9 = (i - 1 + nx) % nx, i1 = (i + 1) % nx;
    const int64_t j9 = (j - 1 + ny) % ny, j1 = (j + 1) % ny;
    const double xterm = c.dxxinv * (a[j * nx + i9] + a[j * nx + i1]);
    const double yterm = c.dyyinv * (a[j9 * nx + i] + a[j1 * nx + i]);
    a[j * nx + i] = (xterm + yterm) * c.dcent;
}

__device__ inline bool late_cell(int64_t j, int64_t i, int64_t nx, int64_t ny)
{
    return ((nx & 1) && i == nx - 1) || ((ny & 1) && j == ny - 1);
}

__global__ void k_gs_color(double* __restrict__ a, const uint8_t* __restrict__ where,
                           int64_t T, int64_t ny, int64_t nx, int parity, GsCoef c)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= T * ny * nx) return;
    const int64_t t = q / (ny * nx), r = q - t * ny * nx;
    const int64_t j = r / nx, i = r - j * nx;
    if (((i + j) & 1) != parity || !where[q] || late_cell(j, i, nx, ny)) return;
    gs_cell(a + t * ny * nx, nx, ny, j, i, c);
}

// number of late cells other than the corner: the last column (odd nx),
// the last row (odd ny); the corner is counted separately when both are odd
__host__ __device__ inline int64_t n_late(int64_t ny, int64_t nx)
{
    const int64_t lx = nx & 1, ly = ny & 1;
    return lx * (ny - ly) + ly * (nx - lx);
}

__device__ inline void late_cell_at(int64_t q, int64_t ny, int64_t nx, int64_t& j, int64_t& i)
{
    const int64_t lx = nx & 1, ly = ny & 1;
    const int64_t ncol = lx * (ny - ly);
    if (q < ncol) {
        j = q;
        i = nx - 1;
    } else {
        j = ny - 1;
        i = q - ncol;
    }
}

// steps 2 and 3 for one slice per workgroup
__global__ void k_gs_late(double* __restrict__ a, const uint8_t* __restrict__ where,
                          int64_t ny, int64_t nx, int parity, GsCoef c)
{
    double* at = a + blockIdx.x * ny * nx;
    const uint8_t* wt = where + blockIdx.x * ny * nx;
    const int64_t nl = n_late(ny, nx);
    for (int64_t q = threadIdx.x; q < nl; q += blockDim.x) {
        int64_t j, i;
        late_cell_at(q, ny, nx, j, i);
        if (((i + j) & 1) == parity && wt[j * nx + i]) gs_cell(at, nx, ny, j, i, c);
    }
    if ((nx & 1) && (ny & 1)) {
        __syncthreads();
        const int64_t j = ny - 1, i = nx - 1;
        if (threadIdx.x == 0 && ((i + j) & 1) == parity && wt[j * nx + i])
            gs_cell(at, nx, ny, j, i, c);
    }
}

constexpr int GS_LDS_BYTES = 64 * 1024;

__host__ __device__ inline int64_t gs_lds_bytes(int64_t cells)
{
    return cells * 8 + ((cells + 15) / 16) * 16;
}

// Every iteration of one slice in LDS: steps 1-3 with workgroup barriers
// between them.  The slice and its mask are read once and written once.
__global__ __launch_bounds__(512) void k_gs_lds(double* __restrict__ a,
                                                const uint8_t* __restrict__ where, int ny,
                                                int nx, int64_t num_iterations, GsCoef c)
{
    extern __shared__ double lds[];
    const int cells = ny * nx;
    double* s = lds;
    uint8_t* w = reinterpret_cast<uint8_t*>(lds + cells);
    double* at = a + (int64_t)blockIdx.x * cells;
    const uint8_t* wt = where + (int64_t)blockIdx.x * cells;
    for (int q = threadIdx.x; q < cells; q += blockDim.x) {
        s[q] = at[q];
        w[q] = wt[q];
    }
    __syncthreads();
    const bool lx = nx & 1, ly = ny & 1;
    const int nl = (int)n_late(ny, nx);
    for (int64_t k = 0; k < num_iterations; k++) {
        for (int parity = 0; parity < 2; parity++) {
            for (int q = threadIdx.x; q < cells; q += blockDim.x) {
                const int j = (unsigned)q / (unsigned)nx, i = q - j * nx;
                if (((i + j) & 1) != parity || !w[q] || late_cell(j, i, nx, ny)) continue;
                gs_cell(s, nx, ny, j, i, c);
            }
            __syncthreads();
            if (nl) {
                for (int q = threadIdx.x; q < nl; q += blockDim.x) {
                    int64_t j, i;
                    late_cell_at(q, ny, nx, j, i);
                    if (((i + j) & 1) == parity && w[j * nx + i]) gs_cell(s, nx, ny, j, i, c);
                }
                __syncthreads();
                if (lx && ly) {
                    const int j = ny - 1, i = nx - 1;
                    if (threadIdx.x == 0 && ((i + j) & 1) == parity && w[j * nx + i])
                        gs_cell(s, nx, ny, j, i, c);
                    __syncthreads();
                }
            }
        }
    }
    for (int q = threadIdx.x; q < cells; q += blockDim.x) at[q] = s[q];
}

// _interp_ndarray.py:208-224 _linear_interpolation
__global__ void k_interp_1d(const double* __restrict__ a, int64_t n,
                            const double* __restrict__ x, int64_t m, int64_t D,
                            double* __restrict__ out)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= D * m) return;
    const int64_t d = q / m;
    const double xv = x[q];
    int64_t x0 = (int64_t)floor(xv);
    if (x0 < 0) x0 = 0;
    else if (x0 > n - 2) x0 = n - 2;
    const double a0 = a[d * n + x0], a1 = a[d * n + x0 + 1];
    const double dx = xv - (double)x0;
    out[q] = a0 * (1 - dx) + a1 * dx;
}

// _interp_ndarray.py:227-269 _bilinear_interpolation
__global__ void k_interp_2d(const double* __restrict__ a, int64_t nx, int64_t ny,
                            const double* __restrict__ x, const double* __restrict__ y,
                            int64_t m, int64_t D, double* __restrict__ out)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= D * m) return;
    const int64_t d = q / m;
    const double xv = x[q], yv = y[q];
    int64_t x00 = (int64_t)floor(xv), y00 = (int64_t)floor(yv);
    if (x00 < 0) x00 = 0;
    else if (x00 > nx - 2) x00 = nx - 2;
    if (y00 < 0) y00 = 0;
    else if (y00 > ny - 2) y00 = ny - 2;
    const double* ad = a + d * nx * ny;
    const double a00 = ad[x00 * ny + y00];
    const double a01 = ad[x00 * ny + y00 + 1];
    const double a10 = ad[(x00 + 1) * ny + y00];
    const double a11 = ad[(x00 + 1) * ny + y00 + 1];
    const double dx = xv - (double)x00, dy = yv - (double)y00;
    const double w00 = (1 - dx) * (1 - dy);
    const double w01 = (1 - dx) * dy;
    const double w10 = dx * (1 - dy);
    const double w11 = dx * dy;
    out[q] = (a00 * w00) + (a01 * w01) + (a10 * w10) + (a11 * w11);
}

}  // namespace

}  // namespace rg

using namespace rg;

extern "C" {

int rg_fill_gauss_seidel_2d(double* a, const uint8_t* where, int64_t T, int64_t ny,
                            int64_t nx, int64_t num_iterations, void* stream)
{
    if (T < 0 || ny < 1 || nx < 1 || num_iterations < 0 || T > 0xffffffffLL)
        return fail(RG_ERR_ARGUMENT, "invalid fill shape or iteration count");
    if (T == 0 || num_iterations == 0) return RG_OK;
    if (!a || !where) return fail(RG_ERR_ARGUMENT, "NULL array");
    hipStream_t s = (hipStream_t)stream;
    const GsCoef c = gs_coef(nx, ny);
    const int64_t cells = T * ny * nx;
    if (gs_lds_bytes(ny * nx) <= GS_LDS_BYTES) {
        const int threads = ny * nx <= 2048 ? 256 : 512;
        k_gs_lds<<<(unsigned)T, threads, gs_lds_bytes(ny * nx), s>>>(a, where, (int)ny, (int)nx,
                                                                      num_iterations, c);
    } else {
        const bool late = (nx & 1) || (ny & 1);
        for (int64_t k = 0; k < num_iterations; k++) {
            for (int parity = 0; parity < 2; parity++) {
                k_gs_color<<<grid_for(cells, 256), 256, 0, s>>>(a, where, T, ny, nx, parity, c);
                if (late) k_gs_late<<<(unsigned)T, 256, 0, s>>>(a, where, ny, nx, parity, c);
            }
        }
    }
    RG_HIP(hipGetLastError());
    return RG_OK;
}

int rg_interp_linear_1d(const double* a, int64_t n, const double* x, int64_t m,
                        int64_t D, double* out, void* stream)
{
    if (D < 0 || n < 2 || m < 0) return fail(RG_ERR_ARGUMENT, "invalid interpolation shape");
    if (D == 0 || m == 0) return RG_OK;
    if (!a || !x || !out) return fail(RG_ERR_ARGUMENT, "NULL array");
    k_interp_1d<<<grid_for(D * m, 256), 256, 0, (hipStream_t)stream>>>(a, n, x, m, D, out);
    RG_HIP(hipGetLastError());
    return RG_OK;
}

int rg_interp_linear_2d(const double* a, int64_t nx, int64_t ny, const double* x,
                        const double* y, int64_t m, int64_t D, double* out, void* stream)
{
    if (D < 0 || nx < 2 || ny < 2 || m < 0)
        return fail(RG_ERR_ARGUMENT, "invalid interpolation shape");
    if (D == 0 || m == 0) return RG_OK;
    if (!a || !x || !y || !out) return fail(RG_ERR_ARGUMENT, "NULL array");
    k_interp_2d<<<grid_for(D * m, 256), 256, 0, (hipStream_t)stream>>>(a, nx, ny, x, y, m, D,
                                                                        out);
    RG_HIP(hipGetLastError());
    return RG_OK;
}

}  // extern "C"
