// rg_search.hip -- point location: the 2-D cell search used by the build
// (index_of_point_secant, _weights_conservative_2d/_grids.py:356-463) as a
// standalone vectorised entry point, the 1-D find_indices searches
// (_find_indices_brute.py:24-51, _find_indices_searchsorted.py:24-62) and
// the cell-area entry point.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "rg_internal.hpp"
#include "rg_wave.hpp"

namespace rg {

int launch_area(const double* x, const double* y, int64_t nx, int64_t ny,
                double* out, hipStream_t s);

// One lane per point: the secant iteration of the reference.  Points the
// reference would hand to the exhaustive search are marked for k_cells_brute.
__global__ void k_cells_secant(const double* __restrict__ x,
                               const double* __restrict__ y, int64_t nx,
                               int64_t ny, const double* __restrict__ px,
                               const double* __restrict__ py, int64_t n,
                               int64_t* __restrict__ cell)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= n) return;
    cell[q] = index_secant(x, y, nx, ny, px[q], py[q]);
}

// One wavefront per point: exhaustive lowest-index search for the points
// marked RG_SEARCH_BRUTE; then flat index -> (i, j) or (INT64_MAX, INT64_MAX).
__global__ void k_cells_finish(const double* __restrict__ x,
                               const double* __restrict__ y, int64_t nx,
                               int64_t ny, const double* __restrict__ px,
                               const double* __restrict__ py, int64_t n,
                               const int64_t* __restrict__ cell,
                               int64_t* __restrict__ ij)
{
    const int64_t q = grid_wave();
    const int lane = threadIdx.x & 63;
    if (q >= n) return;
    int64_t c = cell[q];
    if (c == RG_SEARCH_BRUTE) c = wave_brute(x, y, nx, ny, px[q], py[q], lane);
    if (lane == 0) {
        if (c >= 0) {
            ij[2 * q] = c / (ny - 1);
            ij[2 * q + 1] = c % (ny - 1);
        } else {
            ij[2 * q] = INT64_MAX;
            ij[2 * q + 1] = INT64_MAX;
        }
    }
}

// _find_indices_brute.py:42-49: first m with x[m] <= v <= x[m+1].
__global__ void k_find_brute_1d(const double* __restrict__ xin, int64_t m,
                                const double* __restrict__ xout, int64_t n,
                                int64_t batch, int64_t fill,
                                int64_t* __restrict__ res)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= n * batch) return;
    const int64_t d = t / n;
    const double* a = xin + d * m;
    const double v = xout[t];
    int64_t r = fill;
    for (int64_t q = 0; q < m - 1; q++)
        if (a[q] <= v && v <= a[q + 1]) {
            r = q;
            break;
        }
    res[t] = r;
}

// numpy's searchsorted order: NaN sorts after every number.
__device__ inline bool np_less(double a, double b)
{
    return a < b || (b != b && a == a);
}

// _find_indices_searchsorted.py:38-60 with np.searchsorted(side="left").
__global__ void k_find_sorted_1d(const double* __restrict__ xin, int64_t m,
                                 const double* __restrict__ xout, int64_t n,
                                 int64_t batch, int64_t fill,
                                 int64_t* __restrict__ res)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= n * batch) return;
    const int64_t d = t / n;
    const double* a = xin + d * m;
    const double v = xout[t];
    int64_t lo = 0, hi = m;
    while (lo < hi) {
        int64_t mid = lo + (hi - lo) / 2;
        if (np_less(a[mid], v)) lo = mid + 1;
        else hi = mid;
    }
    int64_t r = lo - 1;
    if (v == a[0]) r = 0;
    else if (r < 0) r = fill;
    else if (r > m - 2) r = fill;
    res[t] = r;
}

namespace {

int check_find_cells(const double* x, const double* y, int64_t nx, int64_t ny,
                     const double* px, const double* py, int64_t* cell_ij)
{
    if (!x || !y || !px || !py || !cell_ij)
        return fail(RG_ERR_ARGUMENT, "NULL argument");
    if (nx < 2 || ny < 2)
        return fail(RG_ERR_ARGUMENT, "grid needs at least 2x2 vertices");
    return RG_OK;
}

// The two kernels of the cell search; `cell` holds one int64 per point
// between them (every entry written by the first before the second reads it).
int launch_find_cells(const double* x, const double* y, int64_t nx, int64_t ny,
                      const double* px, const double* py, int64_t n,
                      int64_t* cell, int64_t* cell_ij, hipStream_t s)
{
    k_cells_secant<<<grid_for(n, 128), 128, 0, s>>>(x, y, nx, ny, px, py, n, cell);
    k_cells_finish<<<grid_for(n * 64, 256), 256, 0, s>>>(x, y, nx, ny, px, py, n,
                                                         cell, cell_ij);
    RG_HIP(hipGetLastError());
    return RG_OK;
}

}  // namespace

// ---- the cell search without allocation -------------------------------------
//
// rg_find_cells_2d_noalloc keeps its per-point scratch in this fixed array
// (one copy per GPU, RG_FIND_CELLS_NOALLOC_MAX * 8 bytes = 16 MiB, held
// until the process ends).  Host code never passes the symbol itself to a
// kernel or copy: on the host it names a host-side stand-in, not device
// memory.  Its device address is looked up once per device, with that device
// current, and kept; calls on one GPU share the array, so calls that
// overlap on one GPU race (issue them on one stream).  Nothing in it needs
// clearing: the secant kernel writes every entry the finishing kernel reads.
static __device__ __align__(256) int64_t scratch[RG_FIND_CELLS_NOALLOC_MAX];

namespace {

constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
int64_t* g_scratch_addr[SCRATCH_MAX_DEVICES] = {};

}  // namespace

int find_cells_scratch(int64_t** out)
{
    *out = nullptr;
    int dev = 0;
    RG_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES)
        return fail(RG_ERR_ARGUMENT, "device " + std::to_string(dev) +
                                         " beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(RG_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                        hipGetErrorString(e));
        g_scratch_addr[dev] = (int64_t*)p;
    }
    *out = g_scratch_addr[dev];
    return RG_OK;
}

}  // namespace rg

using namespace rg;

extern "C" {

int rg_cell_area_2d(const double* x, const double* y, int64_t nx, int64_t ny,
                    double* area, void* stream)
{
    if (nx < 2 || ny < 2) return RG_OK;  // no cells
    if (!x || !y || !area) return fail(RG_ERR_ARGUMENT, "NULL argument");
    return launch_area(x, y, nx, ny, area, (hipStream_t)stream);
}

int rg_find_cells_2d(const double* x, const double* y, int64_t nx, int64_t ny,
                     const double* px, const double* py, int64_t n,
                     int64_t* cell_ij, void* stream)
{
    if (n == 0) return RG_OK;
    RG_TRY(check_find_cells(x, y, nx, ny, px, py, cell_ij));
    hipStream_t s = (hipStream_t)stream;
    DevBuf<int64_t> cell;
    RG_TRY(cell.alloc(n, s));
    return launch_find_cells(x, y, nx, ny, px, py, n, cell.p, cell_ij, s);
}

int rg_find_cells_2d_noalloc(const double* x, const double* y, int64_t nx,
                             int64_t ny, const double* px, const double* py,
                             int64_t n, int64_t* cell_ij, void* stream)
{
    if (n == 0) return RG_OK;
    RG_TRY(check_find_cells(x, y, nx, ny, px, py, cell_ij));
    if (n < 0 || n > RG_FIND_CELLS_NOALLOC_MAX)
        return fail(RG_ERR_ARGUMENT,
                    "rg_find_cells_2d_noalloc: " + std::to_string(n) +
                        " points exceed its fixed scratch of " +
                        std::to_string((long long)RG_FIND_CELLS_NOALLOC_MAX) +
                        " points per call; search them in slices");
    int64_t* cell = nullptr;
    RG_TRY(find_cells_scratch(&cell));
    return launch_find_cells(x, y, nx, ny, px, py, n, cell, cell_ij,
                             (hipStream_t)stream);
}

int rg_find_indices_1d(const double* x_input, int64_t m,
                       const double* x_output, int64_t n, int64_t batch,
                       int method, int64_t fill_value, int64_t* result,
                       void* stream)
{
    if (n == 0 || batch == 0) return RG_OK;
    if (!x_input || !x_output || !result)
        return fail(RG_ERR_ARGUMENT, "NULL argument");
    if (m < 1) return fail(RG_ERR_ARGUMENT, "empty input grid");
    hipStream_t s = (hipStream_t)stream;
    const int64_t tot = n * batch;
    if (method == RG_FIND_BRUTE)
        k_find_brute_1d<<<grid_for(tot, 256), 256, 0, s>>>(
            x_input, m, x_output, n, batch, fill_value, result);
    else if (method == RG_FIND_SEARCHSORTED)
        k_find_sorted_1d<<<grid_for(tot, 256), 256, 0, s>>>(
            x_input, m, x_output, n, batch, fill_value, result);
    else
        return fail(RG_ERR_ARGUMENT, "unknown method");
    RG_HIP(hipGetLastError());
    return RG_OK;
}

}  // extern "C"
