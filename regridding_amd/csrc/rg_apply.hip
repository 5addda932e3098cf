// rg_apply.hip -- application of device-resident weights (the hot loop of
// _regrid/_regrid_from_weights.py:172-182) and its adjoint.
//
// The reference scatters out[i_out] += w * in[i_in] over the triples in
// storage order (sorted by (input, output)).  Each output therefore receives
// its terms in ascending input order, one rounded multiply and one rounded
// add at a time, starting from +0.0.  The GPU gathers instead: one lane owns
// one output cell, walks its CSR row (columns ascending) and accumulates in
// exactly that order, which reproduces the reference bit for bit without
// atomics.  A lane keeps DB batch accumulators so every weight is loaded once
// per DB images; batches larger than DB loop over tiles (the row comes from
// cache the second time).
//
// The adjoint (transpose=1) is the same gather over the CSR by input, whose
// rows hold the outputs in ascending order: exactly the order in which the
// reference applies regrid_from_weights(*transpose_weights(w)).
//
// Launch layout: 256 rows per block; blocks are renumbered so that each of
// the 8 XCDs of the MI355X (which receive blocks round-robin) works on one
// contiguous range of rows, keeping the gathered input neighbourhood in that
// XCD's L2.
#include <hip/hip_runtime.h>

#include <mutex>

#include "rg_internal.hpp"

namespace rg {

constexpr int APPLY_BLOCK = 256;
constexpr int NUM_XCD = 8;

template <class T, int DB>
__global__ __launch_bounds__(APPLY_BLOCK) void k_apply(
    const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
    const T* __restrict__ w, const T* __restrict__ x, T* __restrict__ y,
    int64_t nrows, int64_t ncols, int64_t batch, int64_t ntiles)
{
    const int64_t nb = gridDim.x;  // a multiple of NUM_XCD
    const int64_t b = blockIdx.x;
    const int64_t tile = (b % NUM_XCD) * (nb / NUM_XCD) + b / NUM_XCD;
    if (tile >= ntiles) return;
    const int64_t r = tile * APPLY_BLOCK + threadIdx.x;
    if (r >= nrows) return;
    const int32_t s = rowptr[r], e = rowptr[r + 1];
    for (int64_t d0 = 0; d0 < batch; d0 += DB) {
        const int64_t nd = batch - d0 < DB ? batch - d0 : DB;
        T acc[DB];
#pragma unroll
        for (int j = 0; j < DB; j++) acc[j] = T(0);
        const T* xb = x + d0 * ncols;
        for (int32_t q = s; q < e; q++) {
            const int64_t c = col[q];
            const T wq = w[q];
#pragma unroll
            for (int j = 0; j < DB; j++)
                if (j < nd) acc[j] = acc[j] + wq * xb[j * ncols + c];
        }
        T* yb = y + d0 * nrows + r;
#pragma unroll
        for (int j = 0; j < DB; j++)
            if (j < nd) yb[j * nrows] = acc[j];
    }
}

__global__ void k_to_f32(const double* __restrict__ a, float* __restrict__ b,
                         int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        b[i] = (float)a[i];
}

namespace {

template <class T>
int launch_apply(const int32_t* rowptr, const int32_t* col, const T* w,
                 const T* x, T* y, int64_t nrows, int64_t ncols, int64_t batch,
                 hipStream_t s)
{
    if (nrows == 0 || batch == 0) return RG_OK;
    const int64_t ntiles = (nrows + APPLY_BLOCK - 1) / APPLY_BLOCK;
    const int64_t nb = (ntiles + NUM_XCD - 1) / NUM_XCD * NUM_XCD;
    dim3 g((unsigned)nb), bl(APPLY_BLOCK);
    if (batch >= 16)
        k_apply<T, 16><<<g, bl, 0, s>>>(rowptr, col, w, x, y, nrows, ncols, batch, ntiles);
    else if (batch >= 8)
        k_apply<T, 8><<<g, bl, 0, s>>>(rowptr, col, w, x, y, nrows, ncols, batch, ntiles);
    else if (batch >= 4)
        k_apply<T, 4><<<g, bl, 0, s>>>(rowptr, col, w, x, y, nrows, ncols, batch, ntiles);
    else if (batch >= 2)
        k_apply<T, 2><<<g, bl, 0, s>>>(rowptr, col, w, x, y, nrows, ncols, batch, ntiles);
    else
        k_apply<T, 1><<<g, bl, 0, s>>>(rowptr, col, w, x, y, nrows, ncols, batch, ntiles);
    RG_HIP(hipGetLastError());
    return RG_OK;
}

std::mutex g_f32_mu;

int ensure_f32(rg_weights* W, hipStream_t s)
{
    std::lock_guard<std::mutex> lk(g_f32_mu);
    if (W->val_out_f32 || W->nnz == 0) return RG_OK;
    float *a = nullptr, *b = nullptr;
    RG_TRY(dev_alloc((void**)&a, W->nnz * sizeof(float), s));
    int st = dev_alloc((void**)&b, W->nnz * sizeof(float), s);
    if (st != RG_OK) {
        dev_free(a);
        return st;
    }
    unsigned gs = grid_for(W->nnz, 256);
    if (gs > 8192) gs = 8192;
    k_to_f32<<<gs, 256, 0, s>>>(W->val_in, a, W->nnz);
    k_to_f32<<<gs, 256, 0, s>>>(W->val_out, b, W->nnz);
    RG_HIP(hipGetLastError());
    W->val_in_f32 = a;
    W->val_out_f32 = b;
    return RG_OK;
}

int check_apply(rg_weights* W, const void* vin, void* vout, int64_t batch)
{
    if (!W) return fail(RG_ERR_ARGUMENT, "weights handle is NULL");
    if (batch < 0) return fail(RG_ERR_ARGUMENT, "negative batch");
    if (batch > 0 && (!vin || !vout))
        return fail(RG_ERR_ARGUMENT, "values pointer is NULL");
    if (injected_failure("apply"))
        return fail(RG_ERR_INJECTED, "injected apply failure");
    return RG_OK;
}

}  // namespace

}  // namespace rg

using namespace rg;

extern "C" {

int rg_apply_f64(rg_weights_t w, int transpose, const double* values_input,
                 double* values_output, int64_t batch, void* stream)
{
    RG_TRY(check_apply(w, values_input, values_output, batch));
    hipStream_t s = (hipStream_t)stream;
    if (transpose)
        return launch_apply<double>(w->rowptr_in, w->col_in, w->val_in,
                                    values_input, values_output, w->n_in,
                                    w->n_out, batch, s);
    return launch_apply<double>(w->rowptr_out, w->col_out, w->val_out,
                                values_input, values_output, w->n_out, w->n_in,
                                batch, s);
}

int rg_apply_f32(rg_weights_t w, int transpose, const float* values_input,
                 float* values_output, int64_t batch, void* stream)
{
    RG_TRY(check_apply(w, values_input, values_output, batch));
    hipStream_t s = (hipStream_t)stream;
    RG_TRY(ensure_f32(w, s));
    if (transpose)
        return launch_apply<float>(w->rowptr_in, w->col_in, w->val_in_f32,
                                   values_input, values_output, w->n_in,
                                   w->n_out, batch, s);
    return launch_apply<float>(w->rowptr_out, w->col_out, w->val_out_f32,
                               values_input, values_output, w->n_out, w->n_in,
                               batch, s);
}

}  // extern "C"
