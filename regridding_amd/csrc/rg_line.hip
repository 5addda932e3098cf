// rg_line.hip -- weights along one axis: conservative 1-D and multilinear 1-D.
//
// Conservative 1-D (regridding/_weights/_weights_conservative_1d/
// _weights_conservative_1d.py:57-318): the sweep of the input vertices across
// the output vertices, one element (one line of the orthogonal stack) per
// thread -- the reference's own decomposition (a thread pool over elements,
// _weights_conservative.py:57-102) -- replaying its state machine exactly,
// including the reversal of descending grids, the wraparound (negative)
// indices it emits for them and its first-point search.  A counting pass
// gives every (element, input cell) its number of triples; after a scan the
// emitting pass writes each triple straight to its position in the
// coalesced order (sorted by (input, output); every pair occurs once).
//
// Multilinear 1-D (regridding/_weights/_weights_multilinear.py:148-206): two
// weights per output vertex from the searchsorted index, clamped and marked
// outside as in :95-118, in the reference's storage order.
#include <hip/hip_runtime.h>

#include "rg_internal.hpp"

namespace rg {

namespace {

// The grid of one element viewed ascending (the reference reverses
// descending grids, :88-102).
struct Axis {
    const double* x;
    int64_t n;
    bool rev;
    __device__ double operator()(int64_t k) const { return rev ? x[n - 1 - k] : x[k]; }
};

// _grids.py:40-73 index_of_point: bisection returning the right index.
__device__ int64_t index_of_point(double p, const Axis& g)
{
    int64_t lo = 0, hi = g.n;
    while (hi - lo > 1) {
        const int64_t mid = (lo + hi) / 2;
        if (g(mid) > p) hi = mid;
        else lo = mid;
    }
    return hi;
}

// numpy / numba wraparound of a negative index into an array of length n
__device__ inline int64_t wrap(int64_t i, int64_t n) { return i < 0 ? i + n : i; }

constexpr int64_t OUTSIDE = INT64_MAX;  // sys.maxsize

// The sweep of one element (_weights_conservative_1d.py:78-190).  Calls
// sink(index_input, index_output, ratio, piece) for every triple in the
// reference's order; `piece` counts the triples of the current input cell.
template <class Sink>
__device__ void sweep_1d(const double* xin, int64_t n, const double* xout, int64_t m,
                         const double* wi, Sink& sink)
{
    const Axis xs{xin, n, !(xin[0] < xin[n - 1])};
    const Axis xt{xout, m, !(xout[0] < xout[m - 1])};
    int64_t index_sweep = 0, k;
    double p1 = xs(0);
    bool outside;
    const double t_left = xt(0), t_right = xt(m - 1);
    if (t_left == p1) {
        outside = false;
        k = 0;
    } else if (t_left < p1 && p1 < t_right) {
        outside = false;
        k = index_of_point(p1, xt) - 1;
    } else {
        outside = true;
        k = OUTSIDE;
    }
    int64_t piece = 0, last_cell = -1;
    while (index_sweep < n - 1) {
        double p2 = xs(index_sweep + 1);
        if (outside) {  // _step_outside_static, :193-240
            const double ps = xt(0);
            if (p1 < ps && ps < p2) {
                k = 0;
                p2 = ps;
            } else if (ps == p2) {
                index_sweep++;
                k = 0;
            } else {
                index_sweep++;
            }
            if (k < OUTSIDE) outside = false;
        } else {  // _step_inside_static, :243-318
            int64_t index_input = xs.rev ? ~index_sweep : index_sweep;
            int64_t index_output = xt.rev ? ~k : k;
            const double ps = xt(k + 1);
            if (p1 < ps && ps < p2) {
                k++;
                p2 = ps;
            } else if (ps == p2) {
                k++;
                index_sweep++;
            } else {
                index_sweep++;
            }
            const double length = p2 - p1;
            // length_input comes from the ascending view (:112), indexed
            // with the emitted (possibly negative) input index
            const int64_t li = wrap(index_input, n - 1);
            double ratio = length / (xs(li + 1) - xs(li));
            if (wi) ratio = ratio * wi[li];
            if (li != last_cell) {
                last_cell = li;
                piece = 0;
            }
            sink(index_input, index_output, ratio, li, piece++);
            if (!(0 <= k && k < m - 1)) break;
        }
        p1 = p2;
    }
}

struct CountSink {
    int32_t* cnt;  // per input cell of the element
    __device__ void operator()(int64_t, int64_t, double, int64_t li, int64_t)
    {
        cnt[li]++;
    }
};

struct EmitSink {
    const int64_t* off;  // per input cell: first position (coalesced order)
    const int32_t* cnt;
    bool rev_out;
    int64_t* ii;
    int64_t* io;
    double* w;
    __device__ void operator()(int64_t in, int64_t out, double v, int64_t li,
                               int64_t piece)
    {
        // within an input cell the output index falls when the output grid
        // is descending (the emitted index is ~k): reverse the cell's run
        const int64_t p = off[li] + (rev_out ? cnt[li] - 1 - piece : piece);
        ii[p] = in;
        io[p] = out;
        w[p] = v;
    }
};

__global__ void k_cons1d_count(const double* __restrict__ xin, int64_t n,
                               const double* __restrict__ xout, int64_t m,
                               int64_t D, int32_t* __restrict__ cnt)
{
    const int64_t d = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (d >= D) return;
    int32_t* c = cnt + d * (n - 1);
    for (int64_t i = 0; i < n - 1; i++) c[i] = 0;
    CountSink sink{c};
    sweep_1d(xin + d * n, n, xout + d * m, m, nullptr, sink);
}

__global__ void k_cons1d_emit(const double* __restrict__ xin, int64_t n,
                              const double* __restrict__ xout, int64_t m,
                              int64_t D, const double* __restrict__ wi,
                              const int64_t* __restrict__ off,
                              const int32_t* __restrict__ cnt,
                              int64_t* __restrict__ ii, int64_t* __restrict__ io,
                              double* __restrict__ w)
{
    const int64_t d = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (d >= D) return;
    const double* xo = xout + d * m;
    EmitSink sink{off + d * (n - 1), cnt + d * (n - 1), !(xo[0] < xo[m - 1]), ii, io, w};
    sweep_1d(xin + d * n, n, xo, m, wi ? wi + d * (n - 1) : nullptr, sink);
}

__global__ void k_multilinear_1d(const double* __restrict__ xin, int64_t n,
                                 const double* __restrict__ xout, int64_t m,
                                 int64_t D, const int64_t* __restrict__ index,
                                 const double* __restrict__ wi, int32_t bounds_nan,
                                 int64_t* __restrict__ ii, int64_t* __restrict__ io,
                                 double* __restrict__ w,
                                 unsigned long long* __restrict__ n_outside)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= D * m) return;
    const int64_t d = t / m, i = t - d * m;
    const double* x_in = xin + d * n;
    const double x = xout[t];
    const int64_t index_max = n - 2;
    const int64_t raw = index[t];
    const bool outside = raw < 0 || raw > index_max;
    if (outside) atomicAdd(n_outside, 1ull);
    const bool below = x < x_in[0];
    const int64_t i0 = below ? 0 : (raw < 0 ? 0 : (raw > index_max ? index_max : raw));
    const int64_t i1 = i0 + 1;
    const double x0 = x_in[i0], x1 = x_in[i1];
    double w1 = (x - x0) / (x1 - x0);
    double w0 = 1.0 - w1;
    if (wi) {
        w0 = w0 * wi[d * n + i0];
        w1 = w1 * wi[d * n + i1];
    }
    if (outside && bounds_nan) {
        w0 = __builtin_nan("");
        w1 = __builtin_nan("");
    }
    const int64_t base = d * 2 * m + 2 * i;
    ii[base] = i0;
    io[base] = i;
    w[base] = w0;
    ii[base + 1] = i1;
    io[base + 1] = i;
    w[base + 1] = w1;
}

// _weights_conservative_1d/_grids.py:11-37 cell_length, per element
__global__ void k_cell_length(const double* __restrict__ x, int64_t n, int64_t D,
                              double* __restrict__ out)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= D * (n - 1)) return;
    const int64_t d = t / (n - 1), i = t - d * (n - 1);
    out[t] = x[d * n + i + 1] - x[d * n + i];
}

int check_line(const double* xin, int64_t n, const double* xout, int64_t m, int64_t D)
{
    if (D < 0) return fail(RG_ERR_ARGUMENT, "negative element count");
    if (D > 0 && (!xin || !xout)) return fail(RG_ERR_ARGUMENT, "coordinate pointer is NULL");
    if (n < 2 || m < 2) return fail(RG_ERR_ARGUMENT, "each grid needs at least 2 vertices");
    return RG_OK;
}

}  // namespace

}  // namespace rg

using namespace rg;

extern "C" {

int rg_conservative_1d_count(const double* x_input, int64_t n_input,
                             const double* x_output, int64_t n_output, int64_t D,
                             int32_t* counts, int64_t* offsets, int64_t* total,
                             void* stream)
{
    RG_TRY(check_line(x_input, n_input, x_output, n_output, D));
    if (!counts || !offsets || !total) return fail(RG_ERR_ARGUMENT, "NULL output");
    hipStream_t s = (hipStream_t)stream;
    *total = 0;
    const int64_t cells = D * (n_input - 1);
    if (cells == 0) return RG_OK;
    k_cons1d_count<<<grid_for(D, 64), 64, 0, s>>>(x_input, n_input, x_output,
                                                   n_output, D, counts);
    RG_HIP(hipGetLastError());
    RG_TRY(scan_i32_to_i64(counts, offsets, cells, s));
    RG_TRY(read_scalars(offsets + cells, total, sizeof(int64_t), s));
    return RG_OK;
}

int rg_conservative_1d_emit(const double* x_input, int64_t n_input,
                            const double* x_output, int64_t n_output, int64_t D,
                            const double* weights_input, const int32_t* counts,
                            const int64_t* offsets, int64_t* indices_input,
                            int64_t* indices_output, double* values, void* stream)
{
    RG_TRY(check_line(x_input, n_input, x_output, n_output, D));
    if (D == 0) return RG_OK;
    if (!counts || !offsets || !indices_input || !indices_output || !values)
        return fail(RG_ERR_ARGUMENT, "NULL array");
    hipStream_t s = (hipStream_t)stream;
    k_cons1d_emit<<<grid_for(D, 64), 64, 0, s>>>(x_input, n_input, x_output, n_output,
                                                  D, weights_input, offsets, counts,
                                                  indices_input, indices_output, values);
    RG_HIP(hipGetLastError());
    return RG_OK;
}

int rg_cell_length_1d(const double* x, int64_t n, int64_t D, double* length,
                      void* stream)
{
    if (D < 0 || n < 2) return fail(RG_ERR_ARGUMENT, "need D >= 0 elements of >= 2 vertices");
    if (D == 0) return RG_OK;
    if (!x || !length) return fail(RG_ERR_ARGUMENT, "NULL array");
    hipStream_t s = (hipStream_t)stream;
    k_cell_length<<<grid_for(D * (n - 1), 256), 256, 0, s>>>(x, n, D, length);
    RG_HIP(hipGetLastError());
    return RG_OK;
}

int rg_multilinear_1d(const double* x_input, int64_t n_input, const double* x_output,
                      int64_t n_output, int64_t D, const int64_t* indices,
                      const double* weights_input, int32_t bounds_nan,
                      int64_t* indices_input, int64_t* indices_output, double* values,
                      int64_t* n_outside, void* stream)
{
    RG_TRY(check_line(x_input, n_input, x_output, n_output, D));
    if (!n_outside) return fail(RG_ERR_ARGUMENT, "n_outside is NULL");
    *n_outside = 0;
    const int64_t total = D * n_output;
    if (total == 0) return RG_OK;
    if (!indices || !indices_input || !indices_output || !values)
        return fail(RG_ERR_ARGUMENT, "NULL array");
    hipStream_t s = (hipStream_t)stream;
    DevBuf<unsigned long long> cnt;
    RG_TRY(cnt.alloc(1, s));
    RG_HIP(hipMemsetAsync(cnt.p, 0, sizeof(unsigned long long), s));
    k_multilinear_1d<<<grid_for(total, 256), 256, 0, s>>>(
        x_input, n_input, x_output, n_output, D, indices, weights_input, bounds_nan,
        indices_input, indices_output, values, cnt.p);
    RG_HIP(hipGetLastError());
    unsigned long long h = 0;
    RG_TRY(read_scalars(cnt.p, &h, sizeof(h), s));
    *n_outside = (int64_t)h;
    return RG_OK;
}

}  // extern "C"
