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
// Launch layout.  The rows of a CSR are the cells of a 2-D grid.  A block
// of 256 lanes owns a tile of output cells, tile_w wide along the grid's
// last axis and 256 / tile_w tall, so the input cells it gathers form one
// compact patch (a linear run of rows would be a thin band tilted across
// the input grid, re-fetched by every neighbouring run).  Blocks are
// renumbered so each of the 8 XCDs of the MI355X (blocks arrive
// round-robin) sweeps one contiguous band of tiles, keeping the patches of
// neighbouring tiles in that XCD's L2.  The launch's second dimension is
// the image group: all tiles of images [0, DB) first, then the next DB, so
// the gathered footprint per pass is DB images deep.
//
// The results are stored nontemporal: each is written once and not read
// back by the launch, so written through the caches they would only evict
// the input patches that neighbouring tiles share (at 2049^2 -> 1537^2 and
// 16 images: FP64 0.210 -> 0.205 ms, FP32 0.130 -> 0.123 ms, FP64 adjoint
// 0.256 -> 0.248 ms, FP32 adjoint 0.148 -> 0.094 ms).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "rg_internal.hpp"

namespace rg {

constexpr int APPLY_BLOCK = 256;
constexpr int NUM_XCD = 8;

struct ApplyGrid {
    int64_t gx, gy;       // rows of the CSR as a gx x gy grid of cells
    int64_t tw, th;       // tile shape (tw * th = APPLY_BLOCK)
    int64_t ntx, ntiles;  // tiles along gy, tiles in total
};

template <class T, int DB>
__global__ __launch_bounds__(APPLY_BLOCK) void k_apply(
    const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
    const T* __restrict__ w, const T* __restrict__ x, T* __restrict__ y,
    int64_t nrows, int64_t ncols, int64_t batch, ApplyGrid G)
{
    const int64_t nb = gridDim.x;  // a multiple of NUM_XCD
    const int64_t b = blockIdx.x;
    const int64_t tile = (b % NUM_XCD) * (nb / NUM_XCD) + b / NUM_XCD;
    if (tile >= G.ntiles) return;
    const int64_t ti = tile / G.ntx, tj = tile % G.ntx;
    const int64_t i = ti * G.th + threadIdx.x / G.tw;
    const int64_t j = tj * G.tw + threadIdx.x % G.tw;
    if (i >= G.gx || j >= G.gy) return;
    const int64_t r = i * G.gy + j;
    const int64_t d0 = (int64_t)blockIdx.y * DB;
    const int64_t nd = batch - d0 < DB ? batch - d0 : DB;
    const int32_t s = rowptr[r], e = rowptr[r + 1];
    T acc[DB];
#pragma unroll
    for (int q = 0; q < DB; q++) acc[q] = T(0);
    const T* xb = x + d0 * ncols;
    for (int32_t q = s; q < e; q++) {
        const int64_t c = col[q];
        const T wq = w[q];
#pragma unroll
        for (int m = 0; m < DB; m++)
            if (m < nd) acc[m] = acc[m] + wq * xb[m * ncols + c];
    }
    T* yb = y + d0 * nrows + r;
#pragma unroll
    for (int m = 0; m < DB; m++)
        if (m < nd) __builtin_nontemporal_store(acc[m], &yb[m * nrows]);
}

// ---- the apply plan -------------------------------------------------------
//
// With many images per pass, one tile's gathered footprint (patch cells x
// images x 8 B) outgrows the L1, and the tiles in flight on an XCD outgrow
// its L2: a cache line of the input is evicted before the neighbouring cells
// it also holds are used, and comes back from HBM several times.  The plan
// removes the dependence on cache luck.  Made once per handle and direction
// (the weights are static; every later apply reuses it), it stores for every
// tile the bounding box (i0, j0, rows, cols) of the input cells its rows
// gather, and for every CSR entry the entry's offset inside that box.  The
// planned apply then stages the box of DB images in LDS with one coalesced
// row load per box row, and gathers from LDS.

constexpr int LDS_BYTES = 64 * 1024;  // staged patch per block
constexpr int64_t PLAN_MIN_ROWS = 65536;

// a / b from a float reciprocal and one correction step (the float quotient
// is off by at most one).  Exact for 0 <= a <= 65536 and 1 <= b <= 16384,
// checked exhaustively; a staged patch holds at most 16384 cells.
__device__ inline int32_t fast_div(int32_t a, int32_t b, float inv_b)
{
    int32_t q = (int32_t)((float)a * inv_b);
    const int32_t r = a - q * b;
    if (r < 0) q--;
    else if (r >= b) q++;
    return q;
}

// One input cell of a staged patch: its values in the DB images of the pass,
// side by side.  The patch is cell-major, so a gather reads one entry's DB
// values with one or a few wide LDS loads at one address; an image-major
// patch needs a load and an address computation per image (the FP32 apply
// was bound by that arithmetic).  Images past the pass's count hold 0.
template <class T, int DB>
struct alignas(DB * sizeof(T) >= 16 ? 16 : DB * sizeof(T)) CellVals {
    T v[DB];
};

// Cells each thread stages per step (their DB loads in flight together).
template <class T, int DB>
constexpr int stage_cells_per_step()
{
    return DB >= 16 ? 1 : 16 / DB;
}

// Stages the (rows x cols) box of nd images at xb (image stride ncols, row
// stride src_gy) into `patch`, cell (ii, jj) at ii * cols + jj.  Consecutive
// threads take consecutive cells, so each image's loads are coalesced.
template <class T, int DB>
__device__ inline void stage_cells(CellVals<T, DB>* __restrict__ patch,
                                   const T* __restrict__ xb, int64_t ncols, int64_t src_gy,
                                   int32_t rows, int32_t cols, int32_t nd)
{
    constexpr int SU = stage_cells_per_step<T, DB>();
    const int32_t ncell = rows * cols;
    const float inv_cols = 1.0f / (float)cols;
    for (int32_t c0 = threadIdx.x; c0 < ncell; c0 += APPLY_BLOCK * SU) {
        CellVals<T, DB> cv[SU];
#pragma unroll
        for (int u = 0; u < SU; u++) {
            const int32_t c = c0 + u * APPLY_BLOCK;
            if (c < ncell) {
                const int32_t ii = fast_div(c, cols, inv_cols);
                const T* xc = xb + (int64_t)ii * src_gy + (c - ii * cols);
#pragma unroll
                for (int m = 0; m < DB; m++) cv[u].v[m] = m < nd ? xc[m * ncols] : T(0);
            }
        }
#pragma unroll
        for (int u = 0; u < SU; u++) {
            const int32_t c = c0 + u * APPLY_BLOCK;
            if (c < ncell) patch[c] = cv[u];
        }
    }
}

__global__ __launch_bounds__(APPLY_BLOCK) void k_plan(
    const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
    ApplyGrid G, int64_t src_gy, int32_t* __restrict__ box,
    uint16_t* __restrict__ off)
{
    __shared__ int32_t lo_i, hi_i, lo_j, hi_j;
    if (threadIdx.x == 0) {
        lo_i = INT32_MAX; hi_i = INT32_MIN; lo_j = INT32_MAX; hi_j = INT32_MIN;
    }
    __syncthreads();
    const int64_t tile = blockIdx.x;
    const int64_t ti = tile / G.ntx, tj = tile % G.ntx;
    const int64_t i = ti * G.th + threadIdx.x / G.tw;
    const int64_t j = tj * G.tw + threadIdx.x % G.tw;
    const bool valid = i < G.gx && j < G.gy;
    const int64_t r = i * G.gy + j;
    int32_t s = 0, e = 0;
    if (valid) {
        s = rowptr[r];
        e = rowptr[r + 1];
    }
    // The row's first PLAN_E entries are loaded once, all at once, and kept
    // for both passes (a row holds about five at the usual grid ratios);
    // columns and the row length fit 32 bits, so the divisions are 32-bit.
    constexpr int PLAN_E = 8;
    const uint32_t gy = (uint32_t)src_gy;
    uint32_t cr[PLAN_E];
#pragma unroll
    for (int u = 0; u < PLAN_E; u++) cr[u] = s + u < e ? (uint32_t)col[s + u] : 0u;
    int32_t a0 = INT32_MAX, a1 = INT32_MIN, b0 = INT32_MAX, b1 = INT32_MIN;
#pragma unroll
    for (int u = 0; u < PLAN_E; u++)
        if (s + u < e) {
            const int32_t ci = (int32_t)(cr[u] / gy), cj = (int32_t)(cr[u] - (uint32_t)ci * gy);
            a0 = min(a0, ci); a1 = max(a1, ci);
            b0 = min(b0, cj); b1 = max(b1, cj);
        }
    for (int32_t q = s + PLAN_E; q < e; q++) {
        const uint32_t c = (uint32_t)col[q];
        const int32_t ci = (int32_t)(c / gy), cj = (int32_t)(c - (uint32_t)ci * gy);
        a0 = min(a0, ci); a1 = max(a1, ci);
        b0 = min(b0, cj); b1 = max(b1, cj);
    }
    // the wavefront's extremes first (shuffles), then one LDS atomic per
    // wavefront and bound: 256 threads' atomics on the same four words
    // serialize in the LDS and made this kernel (a row without entries holds
    // the neutral values)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        a0 = min(a0, __shfl_xor(a0, o, 64)); a1 = max(a1, __shfl_xor(a1, o, 64));
        b0 = min(b0, __shfl_xor(b0, o, 64)); b1 = max(b1, __shfl_xor(b1, o, 64));
    }
    if ((threadIdx.x & 63) == 0 && a0 <= a1) {
        atomicMin(&lo_i, a0); atomicMax(&hi_i, a1);
        atomicMin(&lo_j, b0); atomicMax(&hi_j, b1);
    }
    __syncthreads();
    const bool empty = lo_i > hi_i;
    const int32_t rows = empty ? 0 : hi_i - lo_i + 1;
    const int32_t cols = empty ? 0 : hi_j - lo_j + 1;
    if (threadIdx.x == 0) {
        box[4 * tile + 0] = empty ? 0 : lo_i;
        box[4 * tile + 1] = empty ? 0 : lo_j;
        box[4 * tile + 2] = rows;
        box[4 * tile + 3] = cols;
    }
    // offsets are only read for patches that fit the LDS (< 2^16 cells)
    const int64_t cells = (int64_t)rows * cols;
#pragma unroll
    for (int u = 0; u < PLAN_E; u++)
        if (s + u < e) {
            const int32_t ci = (int32_t)(cr[u] / gy), cj = (int32_t)(cr[u] - (uint32_t)ci * gy);
            const int64_t o = (int64_t)(ci - lo_i) * cols + (cj - lo_j);
            off[s + u] = cells < 65536 ? (uint16_t)o : (uint16_t)0;
        }
    for (int32_t q = s + PLAN_E; q < e; q++) {
        const uint32_t c = (uint32_t)col[q];
        const int32_t ci = (int32_t)(c / gy), cj = (int32_t)(c - (uint32_t)ci * gy);
        const int64_t o = (int64_t)(ci - lo_i) * cols + (cj - lo_j);
        off[q] = cells < 65536 ? (uint16_t)o : (uint16_t)0;
    }
}

// One block of PLAN_MAX_BLOCK threads: the largest patch over all tiles
// (each thread's share of the tiles is loaded four at a time).
constexpr int PLAN_MAX_BLOCK = 1024;

__global__ __launch_bounds__(PLAN_MAX_BLOCK) void k_plan_max(const int32_t* __restrict__ box,
                                                             int64_t ntiles,
                                                             int64_t* __restrict__ out)
{
    int64_t m = 0;
    for (int64_t t0 = threadIdx.x; t0 < ntiles; t0 += 4 * PLAN_MAX_BLOCK) {
        int64_t c[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int64_t t = t0 + u * PLAN_MAX_BLOCK;
            c[u] = t < ntiles ? (int64_t)box[4 * t + 2] * box[4 * t + 3] : 0;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) m = c[u] > m ? c[u] : m;
    }
    __shared__ int64_t sh[PLAN_MAX_BLOCK];
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int o = PLAN_MAX_BLOCK / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o && sh[threadIdx.x + o] > sh[threadIdx.x])
            sh[threadIdx.x] = sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

// lds_bytes: the dynamic LDS of the launch; a tile whose patch for the
// block's images does not fit gathers from global memory instead.
template <class T, int DB>
__global__ __launch_bounds__(APPLY_BLOCK) void k_apply_planned(
    const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
    const uint16_t* __restrict__ off, const int32_t* __restrict__ box,
    const T* __restrict__ w, const T* __restrict__ x, T* __restrict__ y,
    int64_t nrows, int64_t ncols, int64_t src_gy, int64_t batch, ApplyGrid G,
    int64_t lds_bytes)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    CellVals<T, DB>* patch = reinterpret_cast<CellVals<T, DB>*>(lds_raw);
    const int64_t nb = gridDim.x;  // a multiple of NUM_XCD
    const int64_t b = blockIdx.x;
    const int64_t tile = (b % NUM_XCD) * (nb / NUM_XCD) + b / NUM_XCD;
    if (tile >= G.ntiles) return;  // uniform per block
    const int32_t i0 = box[4 * tile + 0], j0 = box[4 * tile + 1];
    const int32_t rows = box[4 * tile + 2], cols = box[4 * tile + 3];
    const int64_t cells = (int64_t)rows * cols;
    const int64_t d0 = (int64_t)blockIdx.y * DB;
    const int32_t nd = (int32_t)(batch - d0 < DB ? batch - d0 : DB);
    const bool staged = cells > 0 && cells * DB * (int64_t)sizeof(T) <= lds_bytes;
    if (staged) {
        stage_cells<T, DB>(patch, x + d0 * ncols + (int64_t)i0 * src_gy + j0, ncols, src_gy,
                           rows, cols, nd);
        __syncthreads();
    }
    const int32_t ti = (int32_t)(tile / G.ntx), tj = (int32_t)(tile - (int64_t)ti * G.ntx);
    const int32_t tw = (int32_t)G.tw;
    const int64_t i = (int64_t)ti * G.th + (int32_t)threadIdx.x / tw;
    const int64_t j = (int64_t)tj * G.tw + (int32_t)threadIdx.x % tw;
    if (i >= G.gx || j >= G.gy) return;
    const int64_t r = i * G.gy + j;
    const int32_t s = rowptr[r], e = rowptr[r + 1];
    T acc[DB];
#pragma unroll
    for (int m = 0; m < DB; m++) acc[m] = T(0);
    if (staged) {
        int32_t q = s;
        // four entries' loads in flight at once; the terms are still added
        // one at a time in column order
        for (; q + 4 <= e; q += 4) {
            int32_t o[4];
            T wq[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                o[u] = off[q + u];
                wq[u] = w[q + u];
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const CellVals<T, DB> cv = patch[o[u]];
#pragma unroll
                for (int m = 0; m < DB; m++) acc[m] = acc[m] + wq[u] * cv.v[m];
            }
        }
        for (; q < e; q++) {
            const CellVals<T, DB> cv = patch[off[q]];
            const T wq = w[q];
#pragma unroll
            for (int m = 0; m < DB; m++) acc[m] = acc[m] + wq * cv.v[m];
        }
    } else {
        const T* xb = x + d0 * ncols;
        for (int32_t q = s; q < e; q++) {
            const int64_t c = col[q];
            const T wq = w[q];
#pragma unroll
            for (int m = 0; m < DB; m++)
                if (m < nd) acc[m] = acc[m] + wq * xb[m * ncols + c];
        }
    }
    T* yb = y + d0 * nrows + r;
#pragma unroll
    for (int m = 0; m < DB; m++)
        if (m < nd) __builtin_nontemporal_store(acc[m], &yb[m * nrows]);
}

// The planned apply with all images in one block: the lane's CSR entries
// (up to MAXE) are loaded once into registers, then for each group of DB
// images the block stages the input patch in LDS and gathers from it.  The
// entries and row pointers are read once per launch instead of once per
// image group, and the blocks (one per tile) stay resident longer.
template <class T, int DB, int MAXE>
__global__ __launch_bounds__(APPLY_BLOCK) void k_apply_rounds(
    const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
    const uint16_t* __restrict__ off, const int32_t* __restrict__ box,
    const T* __restrict__ w, const T* __restrict__ x, T* __restrict__ y,
    int64_t nrows, int64_t ncols, int64_t src_gy, int64_t batch, ApplyGrid G,
    int64_t lds_bytes)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    CellVals<T, DB>* patch = reinterpret_cast<CellVals<T, DB>*>(lds_raw);
    const int64_t nb = gridDim.x;  // a multiple of NUM_XCD
    const int64_t b = blockIdx.x;
    const int64_t tile = (b % NUM_XCD) * (nb / NUM_XCD) + b / NUM_XCD;
    if (tile >= G.ntiles) return;  // uniform per block
    const int32_t i0 = box[4 * tile + 0], j0 = box[4 * tile + 1];
    const int32_t rows = box[4 * tile + 2], cols = box[4 * tile + 3];
    const int32_t ncell = rows * cols;
    const int32_t ti = (int32_t)(tile / G.ntx), tj = (int32_t)(tile - (int64_t)ti * G.ntx);
    const int32_t tw = (int32_t)G.tw;
    const int64_t i = (int64_t)ti * G.th + (int32_t)threadIdx.x / tw;
    const int64_t j = (int64_t)tj * G.tw + (int32_t)threadIdx.x % tw;
    const bool valid = i < G.gx && j < G.gy;
    const int64_t r = valid ? i * G.gy + j : 0;
    int32_t s = 0, e = 0;
    if (valid) {
        s = rowptr[r];
        e = rowptr[r + 1];
    }
    const int32_t cnt = e - s;
    int32_t ro[MAXE];
    T rw[MAXE];
#pragma unroll
    for (int q = 0; q < MAXE; q++) {
        if (q < cnt) {
            ro[q] = off[s + q];
            rw[q] = w[s + q];
        }
    }
    const bool fits = ncell > 0 && (int64_t)ncell * DB * (int64_t)sizeof(T) <= lds_bytes;
    for (int64_t d0 = 0; d0 < batch; d0 += DB) {
        const int32_t nd = (int32_t)(batch - d0 < DB ? batch - d0 : DB);
        if (fits) {
            stage_cells<T, DB>(patch, x + d0 * ncols + (int64_t)i0 * src_gy + j0, ncols, src_gy,
                               rows, cols, nd);
            __syncthreads();
        }
        if (valid) {
            T acc[DB];
#pragma unroll
            for (int m = 0; m < DB; m++) acc[m] = T(0);
            if (fits) {
                // (the accumulators of images past nd add 0 and are not stored)
#pragma unroll
                for (int q = 0; q < MAXE; q++) {
                    if (q < cnt) {
                        const CellVals<T, DB> cv = patch[ro[q]];
#pragma unroll
                        for (int m = 0; m < DB; m++) acc[m] = acc[m] + rw[q] * cv.v[m];
                    }
                }
                for (int32_t q = s + MAXE; q < e; q++) {
                    const CellVals<T, DB> cv = patch[off[q]];
                    const T wq = w[q];
#pragma unroll
                    for (int m = 0; m < DB; m++) acc[m] = acc[m] + wq * cv.v[m];
                }
            } else {
                const T* xg = x + d0 * ncols;
                for (int32_t q = s; q < e; q++) {
                    const int64_t c = col[q];
                    const T wq = w[q];
#pragma unroll
                    for (int m = 0; m < DB; m++)
                        if (m < nd) acc[m] = acc[m] + wq * xg[m * ncols + c];
                }
            }
            T* yb = y + d0 * nrows + r;
#pragma unroll
            for (int m = 0; m < DB; m++)
                if (m < nd) __builtin_nontemporal_store(acc[m], &yb[m * nrows]);
        }
        if (fits) __syncthreads();  // the next group overwrites the patch
    }
}

__global__ void k_to_f32(const double* __restrict__ a, float* __restrict__ b,
                         int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        b[i] = (float)a[i];
}

// ---- batched apply ---------------------------------------------------------
//
// Many applies, each with its own weights, in one launch.  Jobs are cut into
// blocks of one tile of output rows x JOB_DB images; block b finds its job in
// the jobs' block prefix.  A tile is a compact patch of the output grid (16
// x 16 cells for 8-byte values), so the inputs its lanes gather are one
// compact patch too, where runs of 256 consecutive rows gathered along a
// thin band (the FP64 apply of the 256-pair batch, 1-8 images per pair: 1.71
// -> 1.51 ms).  The accumulation per output is the one of k_apply.
//
// Blocks are renumbered so each of the 8 XCDs takes one contiguous range of
// the blocks (blocks arrive round-robin over the XCDs): neighbouring tiles
// of a job gather overlapping inputs, which then stay in one XCD's L2
// instead of being fetched by all eight.
template <class T>
__global__ __launch_bounds__(APPLY_BLOCK) void k_apply_jobs(
    const JobDesc* __restrict__ jobs, const int64_t* __restrict__ block_prefix,
    int njobs, int64_t nblocks, const int32_t* __restrict__ abort_flag)
{
    if (abort_flag && *abort_flag) return;  // uniform: one word for the grid
    const int64_t ng = gridDim.x;  // a multiple of NUM_XCD
    const int64_t b = (blockIdx.x % NUM_XCD) * (ng / NUM_XCD) + blockIdx.x / NUM_XCD;
    if (b >= nblocks) return;
    const JobDesc& J = jobs[seg_of(block_prefix, njobs, b)];
    const int64_t lt = b - J.block_base;
    const int64_t grp = lt / J.nrt, tile = lt - grp * J.nrt;
    const int64_t ti = tile / J.ntx, tj = tile - ti * J.ntx;
    const int64_t i = ti * J.th + (int32_t)threadIdx.x / J.tw;
    const int64_t j = tj * J.tw + (int32_t)threadIdx.x % J.tw;
    if (j >= J.gy) return;
    const int64_t r = i * J.gy + j;
    if (r >= J.nrows) return;
    const int64_t d0 = grp * JOB_DB;
    const int64_t nd = J.batch - d0 < JOB_DB ? J.batch - d0 : JOB_DB;
    const int32_t* __restrict__ col = J.col;
    const T* __restrict__ w = (const T*)J.w;
    const T* __restrict__ xb = (const T*)J.x + d0 * J.ncols;
    const int64_t ncols = J.ncols;
    const int32_t s = J.rowptr[r], e = J.rowptr[r + 1];
    T acc[JOB_DB];
#pragma unroll
    for (int m = 0; m < JOB_DB; m++) acc[m] = T(0);
    int32_t q = s;
    for (; q + 4 <= e; q += 4) {
        int64_t c[4];
        T wq[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            c[u] = col[q + u];
            wq[u] = w[q + u];
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int m = 0; m < JOB_DB; m++)
                if (m < nd) acc[m] = acc[m] + wq[u] * xb[m * ncols + c[u]];
    }
    for (; q < e; q++) {
        const int64_t c = col[q];
        const T wq = w[q];
#pragma unroll
        for (int m = 0; m < JOB_DB; m++)
            if (m < nd) acc[m] = acc[m] + wq * xb[m * ncols + c];
    }
    T* yb = (T*)J.y + d0 * J.nrows + r;
#pragma unroll
    for (int m = 0; m < JOB_DB; m++)
        if (m < nd) yb[m * J.nrows] = acc[m];
}

namespace {

// Tile width of the FP64 forward apply; 4-byte values and the adjoint use
// tiles twice as wide (up to APPLY_BLOCK).  Tuned on the plain 2049^2 ->
// 1537^2 pair, 16 images: FP64 forward 16 wide, FP32 32 wide (0.149 ->
// 0.129 ms forward, 0.181 -> 0.148 ms adjoint against 16); the FP64 adjoint
// 32 wide (0.252 -> 0.247 ms against 16; its rows are the input cells, and
// its patches short runs of output rows).
int g_tile_w = 16;
int g_images_per_pass = 8;
// staged patch per block above which fewer images are staged per pass: at
// 32 KiB the FP64 forward apply of the plain 2049^2 -> 1537^2 pair stages 4
// images per pass and holds four blocks per CU instead of three (5 % faster);
// FP32 and the adjoints stage 8 either way
int64_t g_lds_limit = 32 * 1024;
bool g_use_plan = true;
bool g_rounds = true;  // one block per tile looping over image groups
std::mutex g_plan_mu;

ApplyGrid make_grid(const int64_t* shape, int64_t nrows, int64_t value_bytes, int dir)
{
    ApplyGrid G;
    G.gx = shape[0];
    G.gy = shape[1];
    if (G.gx * G.gy != nrows) {  // shape unknown: one long row
        G.gx = 1;
        G.gy = nrows;
    }
    G.tw = value_bytes >= 8 && dir == 0 ? g_tile_w : std::min<int64_t>(2 * g_tile_w, APPLY_BLOCK);
    if (G.gx == 1 || G.tw > G.gy) G.tw = APPLY_BLOCK;
    G.th = APPLY_BLOCK / G.tw;
    G.ntx = (G.gy + G.tw - 1) / G.tw;
    G.ntiles = G.ntx * ((G.gx + G.th - 1) / G.th);
    return G;
}

// The plan of one direction for the current tile shape (made on first use
// on stream s; later applies on other streams must be ordered after it).
int ensure_plan(rg_weights* W, int dir, int64_t value_bytes, const int32_t* rowptr,
                const int32_t* col, const ApplyGrid& G, int64_t src_gy,
                hipStream_t s, rg_weights::Plan** out)
{
    *out = nullptr;
    // below PLAN_MIN_ROWS outputs the gathered input stays in L2 anyway and
    // a plan (two kernels and a host wait) would only add latency
    if (!g_use_plan || G.gx == 1 || W->nnz == 0 || G.gx * G.gy < PLAN_MIN_ROWS)
        return RG_OK;
    std::lock_guard<std::mutex> lk(g_plan_mu);
    rg_weights::Plan& P = W->plan[2 * dir + (value_bytes < 8 ? 1 : 0)];
    if (P.tile_w != G.tw || P.ntiles != G.ntiles) {
        dev_free(P.box);
        dev_free(P.off);
        P.box = nullptr;
        P.off = nullptr;
        P.tile_w = 0;
        RG_TRY(dev_alloc((void**)&P.box, 4 * sizeof(int32_t) * G.ntiles, s));
        RG_TRY(dev_alloc((void**)&P.off, sizeof(uint16_t) * W->nnz, s));
        k_plan<<<(unsigned)G.ntiles, APPLY_BLOCK, 0, s>>>(rowptr, col, G, src_gy,
                                                        P.box, P.off);
        DevBuf<int64_t> mx;
        RG_TRY(mx.alloc(1, s));
        k_plan_max<<<1, PLAN_MAX_BLOCK, 0, s>>>(P.box, G.ntiles, mx.p);
        RG_HIP(hipGetLastError());
        RG_TRY(read_scalars(mx.p, &P.max_cells, sizeof(int64_t), s));
        P.tile_w = G.tw;
        P.ntiles = G.ntiles;
    }
    *out = &P;
    return RG_OK;
}

template <class T>
int launch_apply(rg_weights* W, int dir, const int32_t* rowptr,
                 const int32_t* col, const T* w, const T* x, T* y,
                 int64_t nrows, int64_t ncols, const int64_t* shape,
                 int64_t src_gy, int64_t batch, hipStream_t s)
{
    if (nrows == 0 || batch == 0) return RG_OK;
    const ApplyGrid G = make_grid(shape, nrows, sizeof(T), dir);
    rg_weights::Plan* P = nullptr;
    RG_TRY(ensure_plan(W, dir, sizeof(T), rowptr, col, G, src_gy, s, &P));
    const int64_t nb = (G.ntiles + NUM_XCD - 1) / NUM_XCD * NUM_XCD;
    int db = g_images_per_pass;
    while (db > 1 && db > batch) db >>= 1;
    // staged patches: stage fewer images per pass rather than fall back to
    // the direct gather for the largest tiles
    while (P && db > 1 && P->max_cells * db * (int64_t)sizeof(T) > g_lds_limit) db >>= 1;
    dim3 g((unsigned)nb, (unsigned)((batch + db - 1) / db)), bl(APPLY_BLOCK);
    if (P && g_rounds) {
        int64_t lds = P->max_cells * db * (int64_t)sizeof(T);
        lds = (lds + 255) / 256 * 256;
        if (lds > LDS_BYTES) lds = LDS_BYTES;
        dim3 g1((unsigned)nb);
#define RG_ROUNDS(D)                                                                   \
    k_apply_rounds<T, D, 8><<<g1, bl, (size_t)lds, s>>>(                               \
        rowptr, col, P->off, P->box, w, x, y, nrows, ncols, src_gy, batch, G, lds)
        switch (db) {
        case 16: RG_ROUNDS(16); break;
        case 8: RG_ROUNDS(8); break;
        case 4: RG_ROUNDS(4); break;
        case 2: RG_ROUNDS(2); break;
        default: RG_ROUNDS(1); break;
        }
#undef RG_ROUNDS
    } else if (P) {
        int64_t lds = P->max_cells * db * (int64_t)sizeof(T);
        lds = (lds + 255) / 256 * 256;
        if (lds > LDS_BYTES) lds = LDS_BYTES;
#define RG_PLANNED(D)                                                                  \
    k_apply_planned<T, D><<<g, bl, (size_t)lds, s>>>(                                  \
        rowptr, col, P->off, P->box, w, x, y, nrows, ncols, src_gy, batch, G, lds)
        switch (db) {
        case 16: RG_PLANNED(16); break;
        case 8: RG_PLANNED(8); break;
        case 4: RG_PLANNED(4); break;
        case 2: RG_PLANNED(2); break;
        default: RG_PLANNED(1); break;
        }
#undef RG_PLANNED
    } else {
        switch (db) {
        case 16: k_apply<T, 16><<<g, bl, 0, s>>>(rowptr, col, w, x, y, nrows, ncols, batch, G); break;
        case 8: k_apply<T, 8><<<g, bl, 0, s>>>(rowptr, col, w, x, y, nrows, ncols, batch, G); break;
        case 4: k_apply<T, 4><<<g, bl, 0, s>>>(rowptr, col, w, x, y, nrows, ncols, batch, G); break;
        case 2: k_apply<T, 2><<<g, bl, 0, s>>>(rowptr, col, w, x, y, nrows, ncols, batch, G); break;
        default: k_apply<T, 1><<<g, bl, 0, s>>>(rowptr, col, w, x, y, nrows, ncols, batch, G); break;
        }
    }
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
    // other streams may use the copies as soon as they are published
    RG_HIP(hipStreamSynchronize(s));
    W->val_in_f32 = a;
    W->val_out_f32 = b;
    return RG_OK;
}

std::mutex g_use_mu;

}  // namespace

int note_use(rg_weights* W, hipStream_t s)
{
    if (s == W->home) return RG_OK;
    std::lock_guard<std::mutex> lk(g_use_mu);
    hipEvent_t ev = nullptr;
    for (auto& u : W->uses)
        if (u.first == s) ev = u.second;
    if (!ev) {
        RG_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        W->uses.emplace_back(s, ev);
    }
    RG_HIP(hipEventRecord(ev, s));
    return RG_OK;
}

namespace {

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

// Pinned staging of a batched apply's job table, one per host thread; the
// event marks when the previous upload from it has been consumed.
struct JobUpload {
    void* buf = nullptr;
    size_t cap = 0;
    hipEvent_t done = nullptr;
    int device = -1;  // the device `done` belongs to
    bool recorded = false;
};

int launch_jobs_common(const rg_apply_job* jobs, int64_t njobs, bool f32,
                       hipStream_t s)
{
    if (njobs < 0) return fail(RG_ERR_ARGUMENT, "negative job count");
    if (njobs > 0 && !jobs) return fail(RG_ERR_ARGUMENT, "jobs is NULL");
    std::vector<JobDesc> table;
    std::vector<int64_t> prefix;
    table.reserve(njobs);
    int64_t nblocks = 0;
    for (int64_t k = 0; k < njobs; k++) {
        const rg_apply_job& jb = jobs[k];
        rg_weights* W = jb.weights;
        RG_TRY(check_apply(W, jb.values_input, jb.values_output, jb.batch));
        if (f32) RG_TRY(ensure_f32(W, s));
        JobDesc d;
        if (jb.transpose) {
            d.rowptr = W->rowptr_in;
            d.col = W->col_in;
            d.w = f32 ? (const void*)W->val_in_f32 : (const void*)W->val_in;
            d.nrows = W->n_in;
            d.ncols = W->n_out;
        } else {
            d.rowptr = W->rowptr_out;
            d.col = W->col_out;
            d.w = f32 ? (const void*)W->val_out_f32 : (const void*)W->val_out;
            d.nrows = W->n_out;
            d.ncols = W->n_in;
        }
        d.x = jb.values_input;
        d.y = jb.values_output;
        d.batch = jb.batch;
        if (d.nrows == 0 || d.batch == 0) continue;
        const int64_t* shape = jb.transpose ? W->shape_in : W->shape_out;
        job_tiles(d, shape[0], shape[1], f32 ? 4 : 8);
        d.block_base = nblocks;
        prefix.push_back(nblocks);
        nblocks += d.nrt * ((d.batch + JOB_DB - 1) / JOB_DB);
        table.push_back(d);
    }
    if (table.empty()) return RG_OK;
    if (nblocks >= INT32_MAX) return fail(RG_ERR_ARGUMENT, "batched apply too large");
    const size_t tbytes = table.size() * sizeof(JobDesc);
    const size_t pbytes = prefix.size() * sizeof(int64_t);
    const size_t poff = (tbytes + 255) / 256 * 256, bytes = poff + pbytes;
    static thread_local JobUpload up;
    if (up.recorded) RG_HIP(hipEventSynchronize(up.done));
    if (bytes > up.cap) {
        if (up.buf) (void)hipHostFree(up.buf);
        up.buf = nullptr;
        up.cap = 0;
        const size_t sz = std::max<size_t>(bytes, 1 << 16);
        RG_HIP(hipHostMalloc(&up.buf, sz, hipHostMallocDefault));
        up.cap = sz;
    }
    int device = 0;
    RG_HIP(hipGetDevice(&device));
    if (up.done && up.device != device) {  // an event records on its own device only
        (void)hipEventDestroy(up.done);
        up.done = nullptr;
    }
    if (!up.done) {
        RG_HIP(hipEventCreateWithFlags(&up.done, hipEventDisableTiming));
        up.device = device;
    }
    std::memcpy(up.buf, table.data(), tbytes);
    std::memcpy((char*)up.buf + poff, prefix.data(), pbytes);
    DevBuf<unsigned char> dt;
    RG_TRY(dt.alloc(bytes, s));
    RG_HIP(hipMemcpyAsync(dt.p, up.buf, bytes, hipMemcpyHostToDevice, s));
    RG_HIP(hipEventRecord(up.done, s));
    up.recorded = true;
    const JobDesc* d_jobs = (const JobDesc*)dt.p;
    const int64_t* d_prefix = (const int64_t*)(dt.p + poff);
    const unsigned grid = (unsigned)((nblocks + NUM_XCD - 1) / NUM_XCD * NUM_XCD);
    if (f32)
        k_apply_jobs<float><<<grid, APPLY_BLOCK, 0, s>>>(d_jobs, d_prefix, (int)table.size(),
                                                         nblocks, nullptr);
    else
        k_apply_jobs<double><<<grid, APPLY_BLOCK, 0, s>>>(d_jobs, d_prefix, (int)table.size(),
                                                          nblocks, nullptr);
    RG_HIP(hipGetLastError());
    for (int64_t k = 0; k < njobs; k++) RG_TRY(note_use(jobs[k].weights, s));
    return RG_OK;  // dt returns to the stream's pool after the kernel
}

}  // namespace

int apply_jobs_device(const JobDesc* jobs, const int64_t* prefix, int njobs, int64_t nblocks,
                      const int32_t* abort_flag, hipStream_t s)
{
    if (njobs == 0 || nblocks == 0) return RG_OK;
    if (nblocks >= INT32_MAX) return fail(RG_ERR_ARGUMENT, "batched apply too large");
    const unsigned grid = (unsigned)((nblocks + NUM_XCD - 1) / NUM_XCD * NUM_XCD);
    k_apply_jobs<double><<<grid, APPLY_BLOCK, 0, s>>>(jobs, prefix, njobs, nblocks, abort_flag);
    RG_HIP(hipGetLastError());
    return RG_OK;
}

void job_tiles(JobDesc& d, int64_t gx, int64_t gy, int64_t value_bytes)
{
    if (gx * gy != d.nrows) {  // shape unknown: one long row
        gx = 1;
        gy = d.nrows;
    }
    const int64_t tw = value_bytes >= 8 ? 16 : 32;
    d.tw = gx == 1 || gy < tw ? APPLY_BLOCK : (int32_t)tw;
    d.th = APPLY_BLOCK / d.tw;
    d.gy = gy;
    d.ntx = (gy + d.tw - 1) / d.tw;
    d.nrt = d.ntx * ((gx + d.th - 1) / d.th);
}

}  // namespace rg

using namespace rg;

extern "C" {

int rg_apply_f64(rg_weights_t w, int transpose, const double* values_input,
                 double* values_output, int64_t batch, void* stream)
{
    RG_TRY(check_apply(w, values_input, values_output, batch));
    hipStream_t s = (hipStream_t)stream;
    if (transpose)
        RG_TRY(launch_apply<double>(w, 1, w->rowptr_in, w->col_in, w->val_in,
                                    values_input, values_output, w->n_in,
                                    w->n_out, w->shape_in, w->shape_out[1],
                                    batch, s));
    else
        RG_TRY(launch_apply<double>(w, 0, w->rowptr_out, w->col_out, w->val_out,
                                    values_input, values_output, w->n_out, w->n_in,
                                    w->shape_out, w->shape_in[1], batch, s));
    return note_use(w, s);
}

int rg_apply_f32(rg_weights_t w, int transpose, const float* values_input,
                 float* values_output, int64_t batch, void* stream)
{
    RG_TRY(check_apply(w, values_input, values_output, batch));
    hipStream_t s = (hipStream_t)stream;
    RG_TRY(ensure_f32(w, s));
    if (transpose)
        RG_TRY(launch_apply<float>(w, 1, w->rowptr_in, w->col_in, w->val_in_f32,
                                   values_input, values_output, w->n_in,
                                   w->n_out, w->shape_in, w->shape_out[1],
                                   batch, s));
    else
        RG_TRY(launch_apply<float>(w, 0, w->rowptr_out, w->col_out, w->val_out_f32,
                                   values_input, values_output, w->n_out, w->n_in,
                                   w->shape_out, w->shape_in[1], batch, s));
    return note_use(w, s);
}

int rg_apply_batched_f64(const rg_apply_job* jobs, int64_t njobs, void* stream)
{
    return launch_jobs_common(jobs, njobs, false, (hipStream_t)stream);
}

int rg_apply_batched_f32(const rg_apply_job* jobs, int64_t njobs, void* stream)
{
    return launch_jobs_common(jobs, njobs, true, (hipStream_t)stream);
}

int rg_set_apply_tuning(int tile_w, int images_per_pass, int lds_kib, int rounds)
{
    if (rounds >= 0) g_rounds = rounds != 0;
    // negative tile_w: keep the tile width, turn the staged (planned) apply
    // off; the direct-gather kernel computes the same bits
    if (tile_w < 0) {
        g_use_plan = false;
        tile_w = -tile_w;
    } else if (tile_w > 0) {
        g_use_plan = true;
    }
    if (tile_w) {
        if (tile_w < 1 || tile_w > APPLY_BLOCK || APPLY_BLOCK % tile_w)
            return fail(RG_ERR_ARGUMENT, "tile_w must divide 256");
        g_tile_w = tile_w;
    }
    if (images_per_pass) {
        if (images_per_pass != 1 && images_per_pass != 2 && images_per_pass != 4 &&
            images_per_pass != 8 && images_per_pass != 16)
            return fail(RG_ERR_ARGUMENT, "images_per_pass must be 1, 2, 4, 8 or 16");
        g_images_per_pass = images_per_pass;
    }
    if (lds_kib) {
        if (lds_kib < 1 || lds_kib > LDS_BYTES / 1024)
            return fail(RG_ERR_ARGUMENT, "lds_kib must be in 1..64");
        g_lds_limit = (int64_t)lds_kib * 1024;
    }
    return RG_OK;
}

}  // extern "C"
