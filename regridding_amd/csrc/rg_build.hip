// rg_build.hip -- the conservative 2-D weight build on the GPU.
//
// Pipeline for one grid pair (all on one stream):
//   prepare   input cell areas; per static grid: bbox, boundary polygon,
//             boundary edge table in the outside-scan order, edge boxes
//   start     chunk 0 of every row: the reference's row start (wavefront per
//             row); chunk c > 0: a guessed start from the cell search
//   count     every chunk walks, counts its triples, records its end state,
//             adds its triples to the (input cell, pass) histogram and stages
//             up to `cap` (cin, cout, term) triples per walk thread
//   fixup     per row: chunks whose start differs from the predecessor's end
//             are walked again from the true state (histogram corrected,
//             staging rewritten)
//   emit      scans give each chunk its first emission ordinal and each
//             (cell, pass) its bucket; staged triples are scattered into the
//             buckets as (out << 32 | ordinal, term) records; chunks whose
//             triples overflowed the staging walk a third time
//   reduce    per group of 64 input cells (in LDS): divide by the input cell
//             area, sort each bucket, reduce duplicate pairs in
//             np.add.reduceat order -> CSR by input (the reference order)
//   transpose CSR by output (columns ascending) for the forward apply
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <string>

#include "rg_internal.hpp"
#include "rg_sweep.hpp"
#include "rg_wave.hpp"

namespace rg {

// ---- atomics usable from the shared walk code ----------------------------

// Input-cell buckets are split by sweep pass: bucket (cell, pass) at slot
// 4 * cell + pass, so a cell's four pieces are contiguous (one bucket for
// the reduction) while the triples one pass writes for neighbouring cells
// land close together.

// Histogram increments are combined in registers: a walk's consecutive
// triples mostly fall in one or two (cell, pass) buckets (the static cell of
// a sweep over the output grid; the left and right cells of an input edge),
// so two pending (bucket, count) entries replace most atomics.
struct HistCache {
    int64_t b0 = -1, b1 = -1;
    int32_t n0 = 0, n1 = 0;
    __device__ void add(int32_t* __restrict__ hist, int64_t b, int32_t d)
    {
        if (b == b0) {
            n0 += d;
        } else if (b == b1) {
            n1 += d;
        } else {
            if (b1 >= 0 && n1 != 0) atomicAdd(&hist[b1], n1);
            b1 = b0;
            n1 = n0;
            b0 = b;
            n0 = d;
        }
    }
    __device__ void flush(int32_t* __restrict__ hist)
    {
        if (b0 >= 0 && n0 != 0) atomicAdd(&hist[b0], n0);
        if (b1 >= 0 && n1 != 0) atomicAdd(&hist[b1], n1);
        b0 = b1 = -1;
        n0 = n1 = 0;
    }
};

// Counting walk: counts the chunk's triples, adds them to the per-input-cell
// histogram and keeps the first `cap` of them in the staging buffer.  Call
// flush() after the walk.
struct StageSink {
    int32_t n;
    int64_t g, cap;
    Triple* __restrict__ stage;
    int32_t* __restrict__ hist;
    int32_t p;
    HistCache hc;
    __device__ void emit(int64_t cin, int64_t cout, double a)
    {
        if (n < cap) {
            Triple t;
            t.cin = (int32_t)cin;
            t.cout = (int32_t)cout;
            t.a = a;
            stage[stage_index(g, n, cap)] = t;
        }
        n++;
        hc.add(hist, 4 * cin + p, 1);
    }
    __device__ void flush() { hc.flush(hist); }
};

struct UncountSink {
    int32_t* hist;
    int32_t p;
    HistCache hc;
    __device__ void emit(int64_t cin, int64_t, double)
    {
        hc.add(hist, 4 * cin + p, -1);
    }
    __device__ void flush() { hc.flush(hist); }
};

// Bucket cursors: cursor[b] starts at the first record of bucket b and each
// store claims the next one with one returning atomic (record positions fit
// in 32 bits: a build holds fewer than 2^32 triples).  A store is only
// bounds-checked against the record count; k_cursor_check then verifies that
// every cursor ended exactly at its bucket's end, which holds iff each
// bucket received exactly the triples the count walk histogrammed.
__device__ inline void put_records(uint32_t* __restrict__ cursor,
                                   KV* __restrict__ rec, uint32_t n_rec,
                                   int32_t* __restrict__ overflow, int64_t b,
                                   int64_t cout, uint32_t ordinal, double a)
{
    const uint32_t pos = atomicAdd(&cursor[b], 1u);
    if (pos < n_rec) {
        KV r;
        r.key = ((uint64_t)(uint32_t)cout << 32) | (uint64_t)ordinal;
        r.val = a;
        rec[pos] = r;
    } else {
        atomicAdd(overflow, 1);
    }
}

// Writes each triple into its input cell's bucket (chunks whose triples
// overflowed the staging buffer walk again with this sink).
struct BucketSink {
    uint32_t ordinal;
    uint32_t* __restrict__ cursor;
    KV* __restrict__ rec;
    uint32_t n_rec;
    int32_t* __restrict__ overflow;
    int32_t p;
    __device__ void emit(int64_t cin, int64_t cout, double a)
    {
        put_records(cursor, rec, n_rec, overflow, 4 * cin + p, cout, ordinal, a);
        ordinal++;
    }
};

__global__ void k_cursor_init(const int64_t* __restrict__ bucket_off,
                              uint32_t* __restrict__ cursor, int64_t nb)
{
    const int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (b < nb) cursor[b] = (uint32_t)bucket_off[b];
}

__global__ void k_cursor_check(const int64_t* __restrict__ bucket_off,
                               const uint32_t* __restrict__ cursor, int64_t nb,
                               int32_t* __restrict__ bad)
{
    const int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (b < nb && (int64_t)cursor[b] != bucket_off[b + 1]) atomicAdd(bad, 1);
}

// ---- prepare -------------------------------------------------------------

__global__ void k_area(const double* __restrict__ x, const double* __restrict__ y,
                       int64_t nx, int64_t ny, double* __restrict__ out)
{
    const int64_t ncy = ny - 1;
    const int64_t n = (nx - 1) * ncy;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n;
         c += (int64_t)gridDim.x * blockDim.x) {
        out[c] = cell_area(x, y, ny, c / ncy, c % ncy);
    }
}

int launch_area(const double* x, const double* y, int64_t nx, int64_t ny,
                double* out, hipStream_t s)
{
    const int64_t n = (nx - 1) * (ny - 1);
    if (n <= 0) return RG_OK;
    k_area<<<std::min<unsigned>(grid_for(n, 256), 65536), 256, 0, s>>>(x, y, nx,
                                                                       ny, out);
    RG_HIP(hipGetLastError());
    return RG_OK;
}

// numpy min/max semantics: a NaN anywhere makes the result NaN.
__device__ inline double nan_min(double a, double b)
{
    return (a != a || b != b) ? __builtin_nan("") : (a < b ? a : b);
}
__device__ inline double nan_max(double a, double b)
{
    return (a != a || b != b) ? __builtin_nan("") : (a > b ? a : b);
}

constexpr int BBOX_BLOCK = 256;

__global__ void k_bbox_partial(const double* __restrict__ x,
                               const double* __restrict__ y, int64_t n,
                               double* __restrict__ part)
{
    double v[4] = {x[0], y[0], x[0], y[0]};
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        v[0] = nan_min(v[0], x[i]);
        v[1] = nan_min(v[1], y[i]);
        v[2] = nan_max(v[2], x[i]);
        v[3] = nan_max(v[3], y[i]);
    }
    __shared__ double sh[4][BBOX_BLOCK];
    for (int q = 0; q < 4; q++) sh[q][threadIdx.x] = v[q];
    __syncthreads();
    for (int o = BBOX_BLOCK / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            sh[0][threadIdx.x] = nan_min(sh[0][threadIdx.x], sh[0][threadIdx.x + o]);
            sh[1][threadIdx.x] = nan_min(sh[1][threadIdx.x], sh[1][threadIdx.x + o]);
            sh[2][threadIdx.x] = nan_max(sh[2][threadIdx.x], sh[2][threadIdx.x + o]);
            sh[3][threadIdx.x] = nan_max(sh[3][threadIdx.x], sh[3][threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 4; q++) part[4 * blockIdx.x + q] = sh[q][0];
}

__global__ void k_bbox_final(const double* __restrict__ part, int nparts,
                             double* __restrict__ bbox)
{
    double v[4] = {part[0], part[1], part[2], part[3]};
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
        v[0] = nan_min(v[0], part[4 * b + 0]);
        v[1] = nan_min(v[1], part[4 * b + 1]);
        v[2] = nan_max(v[2], part[4 * b + 2]);
        v[3] = nan_max(v[3], part[4 * b + 3]);
    }
    __shared__ double sh[4][BBOX_BLOCK];
    for (int q = 0; q < 4; q++) sh[q][threadIdx.x] = v[q];
    __syncthreads();
    for (int o = BBOX_BLOCK / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            sh[0][threadIdx.x] = nan_min(sh[0][threadIdx.x], sh[0][threadIdx.x + o]);
            sh[1][threadIdx.x] = nan_min(sh[1][threadIdx.x], sh[1][threadIdx.x + o]);
            sh[2][threadIdx.x] = nan_max(sh[2][threadIdx.x], sh[2][threadIdx.x + o]);
            sh[3][threadIdx.x] = nan_max(sh[3][threadIdx.x], sh[3][threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 4; q++) bbox[q] = sh[q][0];
}

__global__ void k_interleave(const double* __restrict__ x,
                             const double* __restrict__ y, int64_t n,
                             double2* __restrict__ xy)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        double2 v;
        v.x = x[i];
        v.y = y[i];
        xy[i] = v;
    }
}

// Boundary polygon and boundary edge table (same length 2(nx-1)+2(ny-1)).
__global__ void k_boundary(const double* __restrict__ x,
                           const double* __restrict__ y, int64_t nx, int64_t ny,
                           double* __restrict__ bpx, double* __restrict__ bpy,
                           double* __restrict__ be, int32_t* __restrict__ bcell)
{
    const int64_t n = 2 * (nx - 1) + 2 * (ny - 1);
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= n) return;
    int64_t v = boundary_vertex(nx, ny, e);
    bpx[e] = x[v];
    bpy[e] = y[v];
    boundary_edge(x, y, nx, ny, e, be + 4 * e, bcell + e);
}

__global__ void k_edge_boxes(const double* __restrict__ be, int64_t nbe,
                             Box* __restrict__ bblk, int64_t nblk)
{
    const int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (b >= nblk) return;
    Box r;
    r.x0 = r.y0 = __builtin_huge_val();
    r.x1 = r.y1 = -__builtin_huge_val();
    int64_t e1 = (b + 1) * BND_BLOCK < nbe ? (b + 1) * BND_BLOCK : nbe;
    for (int64_t e = b * BND_BLOCK; e < e1; e++) {
        const double* q = be + 4 * e;
        // fmin/fmax ignore NaN; an edge with a NaN end never hits anyway
        r.x0 = fmin(r.x0, fmin(q[0], q[2]));
        r.x1 = fmax(r.x1, fmax(q[0], q[2]));
        r.y0 = fmin(r.y0, fmin(q[1], q[3]));
        r.y1 = fmax(r.y1, fmax(q[1], q[3]));
    }
    bblk[b] = r;
}

__global__ void k_super_boxes(const Box* __restrict__ bblk, int64_t nblk,
                              Box* __restrict__ bsup, int64_t nsup)
{
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= nsup) return;
    Box r;
    r.x0 = r.y0 = __builtin_huge_val();
    r.x1 = r.y1 = -__builtin_huge_val();
    int64_t b1 = (s + 1) * BND_SUPER < nblk ? (s + 1) * BND_SUPER : nblk;
    for (int64_t b = s * BND_SUPER; b < b1; b++) {
        r.x0 = fmin(r.x0, bblk[b].x0);
        r.y0 = fmin(r.y0, bblk[b].y0);
        r.x1 = fmax(r.x1, bblk[b].x1);
        r.y1 = fmax(r.y1, bblk[b].y1);
    }
    bsup[s] = r;
}

// ---- start states --------------------------------------------------------

// Row start of the reference (:293-322), one wavefront per row: bounding
// box, winding number over the static boundary split across the lanes,
// secant search on lane 0 with the exhaustive fallback spread over the wave.
__global__ void k_row_start(Build B, int32_t* __restrict__ start,
                            int32_t* __restrict__ anomalies)
{
    const int64_t gw = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (gw >= B.nrows_total) return;  // uniform per wavefront
    const int p = pass_of_row(B, gw);
    const Pass& P = B.pass[p];
    const StaticGrid& s = P.st;
    const int64_t row = gw - B.row_base[p];
    const int64_t g = P.chunk_base + row * P.cpr;
    const double px = sweep_x(P, row, 0), py = sweep_y(P, row, 0);
    int32_t state = STATE_OUTSIDE;
    if (s.bbox[0] <= px && px <= s.bbox[2] && s.bbox[1] <= py &&
        py <= s.bbox[3]) {
        double w = 0;
        for (int64_t k = lane; k < s.nbp; k += 64) {
            int64_t km = k == 0 ? s.nbp - 1 : k - 1;
            w += winding_term(px, py, s.bpx[km], s.bpy[km], s.bpx[k], s.bpy[k]);
        }
        w = wave_sum(w);
        if (w != 0) {
            int64_t c = 0;
            if (lane == 0) c = index_secant(s.x, s.y, s.nx, s.ny, px, py);
            c = __shfl(c, 0, 64);
            if (c == RG_SEARCH_BRUTE) c = wave_brute(s.x, s.y, s.nx, s.ny, px, py, lane);
            if (c >= 0) {
                state = (int32_t)c;
            } else if (lane == 0) {
                // inside the boundary polygon but in no cell: the reference
                // then indexes with sys.maxsize; treated as outside here
                atomicAdd(anomalies, 1);
            }
        }
    }
    if (lane == 0) start[g] = state;
}

__global__ void k_chunk_start(Build B, int32_t* __restrict__ start)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= B.nthreads) return;
    const ChunkRef r = chunk_of_thread(B, t);
    if (!r.valid || r.c == 0) return;
    start[r.g] = chunk_start_guess(B.pass[r.p], r.row, r.c * B.K);
}

// ---- walks ---------------------------------------------------------------

__device__ inline void chunk_range(const Build& B, const Pass& P, int64_t c,
                                   int64_t* k0, int64_t* k1)
{
    *k0 = c * B.K;
    int64_t e = *k0 + B.K;
    *k1 = e < P.nk - 1 ? e : P.nk - 1;
    if (*k1 < *k0) *k1 = *k0;
}

__global__ void k_walk_count(Build B, const int32_t* __restrict__ start,
                             int32_t* __restrict__ end,
                             int32_t* __restrict__ count,
                             int32_t* __restrict__ hist,
                             Triple* __restrict__ stage, int64_t cap)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= B.nthreads) return;
    const ChunkRef r = chunk_of_thread(B, t);
    if (!r.valid) return;
    const Pass& P = B.pass[r.p];
    int64_t k0, k1;
    chunk_range(B, P, r.c, &k0, &k1);
    StageSink sink{0, t, cap, stage, hist, r.p, {}};
    end[r.g] = walk_chunk(B, P, r.row, k0, k1, start[r.g], sink);
    sink.flush();
    count[r.g] = sink.n;
}

// One thread per row: make every chunk start where its predecessor ended.
__global__ void k_fixup(Build B, int32_t* __restrict__ start,
                        int32_t* __restrict__ end, int32_t* __restrict__ count,
                        int32_t* __restrict__ hist, int32_t* __restrict__ nfix,
                        int32_t* __restrict__ failed,
                        Triple* __restrict__ stage, int64_t cap)
{
    const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r >= B.nrows_total) return;
    const int p = pass_of_row(B, r);
    const Pass& P = B.pass[p];
    const int64_t row = r - B.row_base[p];
    const int64_t g0 = P.chunk_base + row * P.cpr;
    int32_t fixes = 0;
    for (int64_t c = 1; c < P.cpr; c++) {
        const int64_t g = g0 + c;
        const int32_t truth = end[g - 1];
        if (truth == STATE_FAILED) break;
        if (start[g] == truth) continue;
        int64_t k0, k1;
        chunk_range(B, P, c, &k0, &k1);
        UncountSink undo{hist, p, {}};
        (void)walk_chunk(B, P, row, k0, k1, start[g], undo);
        undo.flush();
        StageSink redo{0, thread_of_chunk(P, row, c), cap, stage, hist, p, {}};
        end[g] = walk_chunk(B, P, row, k0, k1, truth, redo);
        redo.flush();
        count[g] = redo.n;
        start[g] = truth;
        fixes++;
    }
    for (int64_t c = 0; c < P.cpr; c++)
        if (end[g0 + c] == STATE_FAILED) {
            atomicAdd(failed, 1);
            break;
        }
    if (fixes) atomicAdd(nfix, fixes);
}

// Staged triples -> input-cell buckets, each tagged with its emission
// ordinal (the chunk's first ordinal plus its position in the chunk).
// Chunks whose triples overflowed the staging walk again, emitting straight
// into the buckets.
__global__ void k_scatter(Build B, const int32_t* __restrict__ start,
                          const int32_t* __restrict__ count,
                          const Triple* __restrict__ stage, int64_t cap,
                          const int64_t* __restrict__ chunk_off,
                          uint32_t* __restrict__ cursor,
                          KV* __restrict__ rec, uint32_t n_rec,
                          int32_t* __restrict__ overflow,
                          int32_t* __restrict__ restaged)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= B.nthreads) return;
    const ChunkRef r = chunk_of_thread(B, t);
    if (!r.valid) return;
    const int32_t n = count[r.g];
    const uint32_t ord = (uint32_t)chunk_off[r.g];
    if (n > cap) {
        const Pass& P = B.pass[r.p];
        int64_t k0, k1;
        chunk_range(B, P, r.c, &k0, &k1);
        BucketSink sink{ord, cursor, rec, n_rec, overflow, r.p};
        (void)walk_chunk(B, P, r.row, k0, k1, start[r.g], sink);
        atomicAdd(restaged, 1);
        return;
    }
    // consecutive staged triples of one bucket (a walk inside one static
    // cell of a sweep over the output grid) claim their records with one
    // atomic
    int32_t q0 = 0;
    while (q0 < n) {
        const Triple e0 = stage[stage_index(t, q0, cap)];
        int32_t q1 = q0 + 1;
        while (q1 < n && q1 - q0 < 16 && stage[stage_index(t, q1, cap)].cin == e0.cin)
            q1++;
        const uint32_t m = (uint32_t)(q1 - q0);
        const uint32_t pos = atomicAdd(&cursor[4 * (int64_t)e0.cin + r.p], m);
        if ((uint64_t)pos + m <= n_rec) {
            for (int32_t q = q0; q < q1; q++) {
                const Triple e = q == q0 ? e0 : stage[stage_index(t, q, cap)];
                KV v;
                v.key = ((uint64_t)(uint32_t)e.cout << 32) | (uint64_t)(ord + (uint32_t)q);
                v.val = e.a;
                rec[pos + (uint32_t)(q - q0)] = v;
            }
        } else {
            atomicAdd(overflow, 1);
        }
        q0 = q1;
    }
}

// ---- reduce and CSR ------------------------------------------------------

// Coalescing, one wavefront per group of 64 consecutive input cells (whose
// buckets are one contiguous range).  The range is loaded into LDS with
// coalesced loads, each lane sorts and reduces its own bucket there
// (bucket_reduce), and the wave writes the group's unique (out, weight)
// pairs back, packed from the start of the range, in row order.  A group
// whose range exceeds the LDS share reduces in global memory instead and is
// marked so the CSR pass copies it bucket by bucket.
constexpr int RED_WAVES = 2;     // wavefronts per block
constexpr int RED_CAP = 1536;    // LDS entries per wavefront (24 KiB)

__global__ __launch_bounds__(RED_WAVES * 64) void k_reduce_groups(
    const int64_t* __restrict__ bucket_off, KV* __restrict__ rec,
    const double* __restrict__ vol_in, const double* __restrict__ wi,
    int64_t n_in, int32_t* __restrict__ nnz_row,
    int32_t* __restrict__ group_mode, int32_t* __restrict__ long_list,
    int32_t* __restrict__ n_long)
{
    __shared__ uint64_t sk[RED_WAVES][RED_CAP];
    __shared__ double sv[RED_WAVES][RED_CAP];
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t grp = (int64_t)blockIdx.x * RED_WAVES + wid;
    const int64_t i0 = grp * 64;
    if (i0 >= n_in) return;  // uniform per wavefront; no block barriers below
    const int64_t i1 = i0 + 64 < n_in ? i0 + 64 : n_in;
    const int64_t i = i0 + lane;
    const int64_t base = bucket_off[4 * i0];
    const int64_t total = bucket_off[4 * i1] - base;
    int64_t s = 0, n = 0;
    double vol = 1;
    if (i < i1) {
        s = bucket_off[4 * i] - base;
        n = bucket_off[4 * i + 4] - bucket_off[4 * i];
        vol = vol_in[i];
    }
    bool lr = false;
    if (total <= RED_CAP) {
        uint64_t* K = sk[wid];
        double* V = sv[wid];
        for (int64_t q = lane; q < total; q += 64) {
            const KV r = rec[base + q];
            K[q] = r.key;
            V[q] = r.val;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        int64_t m = 0;
        if (i < i1) {
            for (int64_t q = 0; q < n; q++) V[s + q] = weight_of(V[s + q], vol, wi, i);
            m = bucket_reduce(K + s, V + s, n, &lr);
        }
        if (lr) m = n;  // left sorted, uncompacted, for k_bucket_long
        const bool packed = __ballot(lr) == 0;
        int64_t dst = s;  // a group with a long run keeps bucket starts
        if (packed) {
            // exclusive prefix of m over the wave: packed destination
            int64_t incl = m;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                int64_t y = __shfl_up(incl, o, 64);
                if (lane >= o) incl += y;
            }
            dst = incl - m;
        }
        for (int64_t t = 0; t < m; t++) {
            KV r;
            r.key = K[s + t];
            r.val = V[s + t];
            rec[base + dst + t] = r;
        }
        if (i < i1) {
            nnz_row[i] = (int32_t)m;
            if (lr) long_list[atomicAdd(n_long, 1)] = (int32_t)i;
        }
        if (lane == 0) group_mode[grp] = packed ? 0 : 1;
    } else {
        // oversized group: each lane finishes its bucket in global memory
        if (i < i1) {
            KV* r = rec + base + s;
            for (int64_t q = 0; q < n; q++) r[q].val = weight_of(r[q].val, vol, wi, i);
            sort_records(r, n);
            nnz_row[i] = (int32_t)bucket_reduce_long(r, n);
        }
        if (lane == 0) group_mode[grp] = 1;
    }
}

// Buckets with a run of more than 129 terms (see bucket_reduce): left
// sorted at their bucket start (their group is never packed) and finished
// here.  Grid-stride over the normally empty list; fixes nnz_row.
__global__ void k_bucket_long(const int64_t* __restrict__ bucket_off,
                              KV* __restrict__ rec,
                              const int32_t* __restrict__ long_list,
                              const int32_t* __restrict__ n_long,
                              int32_t* __restrict__ nnz_row)
{
    const int32_t nl = *n_long;
    for (int32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < nl;
         q += gridDim.x * blockDim.x) {
        const int64_t i = long_list[q];
        const int64_t start = bucket_off[4 * i];
        nnz_row[i] = (int32_t)bucket_reduce_long(rec + start,
                                                 bucket_off[4 * i + 4] - start);
    }
}

// CSR by input from the reduced groups: a packed group is one contiguous
// copy; a fallback group copies bucket by bucket.  Also counts the entries
// of every output cell for the transpose.
__global__ void k_csr_groups(const int64_t* __restrict__ bucket_off,
                             const int64_t* __restrict__ row_off,
                             const int32_t* __restrict__ group_mode,
                             const KV* __restrict__ rec, int64_t n_in,
                             int32_t* __restrict__ rowptr,
                             int32_t* __restrict__ col,
                             double* __restrict__ out_val,
                             int32_t* __restrict__ hist_out)
{
    const int64_t grp = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    const int64_t i0 = grp * 64;
    if (i0 > n_in) return;
    if (i0 == n_in) {  // the closing row pointer
        if (lane == 0) rowptr[n_in] = (int32_t)row_off[n_in];
        return;
    }
    const int64_t i1 = i0 + 64 < n_in ? i0 + 64 : n_in;
    if (i0 + lane < i1) rowptr[i0 + lane] = (int32_t)row_off[i0 + lane];
    if (i1 == n_in && lane == 0) rowptr[n_in] = (int32_t)row_off[n_in];
    const int64_t d0 = row_off[i0];
    if (group_mode[grp] == 0) {
        const int64_t b0 = bucket_off[4 * i0], cnt = row_off[i1] - d0;
        for (int64_t q = lane; q < cnt; q += 64) {
            const KV r = rec[b0 + q];
            const int32_t o = (int32_t)r.key;
            col[d0 + q] = o;
            out_val[d0 + q] = r.val;
            atomicAdd(&hist_out[o], 1);
        }
    } else if (i0 + lane < i1) {
        const int64_t i = i0 + lane;
        const int64_t sb = bucket_off[4 * i], d = row_off[i], m = row_off[i + 1] - d;
        for (int64_t q = 0; q < m; q++) {
            const KV r = rec[sb + q];
            const int32_t o = (int32_t)r.key;
            col[d + q] = o;
            out_val[d + q] = r.val;
            atomicAdd(&hist_out[o], 1);
        }
    }
}

// Transpose scatter: entry q of row i of a CSR goes to row col[q] of the
// transposed CSR at an atomically claimed slot; k_rows_sort then orders each
// transposed row by column, which makes the result deterministic.
__global__ void k_transpose_scatter(const int32_t* __restrict__ rowptr,
                                    const int32_t* __restrict__ col,
                                    const double* __restrict__ val,
                                    int64_t nrows,
                                    const int64_t* __restrict__ t_off,
                                    int32_t* __restrict__ cursor,
                                    int32_t* __restrict__ t_col,
                                    double* __restrict__ t_val)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= nrows) return;
    for (int32_t q = rowptr[i]; q < rowptr[i + 1]; q++) {
        int32_t o = col[q];
        int64_t pos = t_off[o] + atomicAdd(&cursor[o], 1);
        if (pos < t_off[o + 1]) {
            t_col[pos] = (int32_t)i;
            t_val[pos] = val[q];
        }
    }
}

// Orders every row of the transposed CSR by column (the rows were filled
// in atomic order).  One wavefront per 64 consecutive rows, whose entries
// are one contiguous range: loaded into LDS with coalesced loads, sorted
// per lane in LDS, stored back coalesced.  Oversized groups sort in global
// memory.  Also writes the int32 row pointers.
constexpr int SORT_WAVES = 4;
constexpr int SORT_CAP = 1536;  // entries per wavefront (18 KiB)

__global__ __launch_bounds__(SORT_WAVES * 64) void k_rows_sort(
    const int64_t* __restrict__ off, int64_t nrows, int32_t* __restrict__ col,
    double* __restrict__ val, int32_t* __restrict__ rowptr)
{
    __shared__ int32_t sc[SORT_WAVES][SORT_CAP];
    __shared__ double sv[SORT_WAVES][SORT_CAP];
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t r0 = ((int64_t)blockIdx.x * SORT_WAVES + wid) * 64;
    if (r0 > nrows) return;  // uniform per wavefront
    const int64_t r = r0 + lane;
    if (r <= nrows) rowptr[r] = (int32_t)off[r];
    if (r0 == nrows) return;
    const int64_t r1 = r0 + 64 < nrows ? r0 + 64 : nrows;
    const int64_t base = off[r0], total = off[r1] - base;
    int64_t s = 0, n = 0;
    if (r < r1) {
        s = off[r] - base;
        n = off[r + 1] - off[r];
    }
    if (total <= SORT_CAP) {
        int32_t* C = sc[wid];
        double* V = sv[wid];
        for (int64_t q = lane; q < total; q += 64) {
            C[q] = col[base + q];
            V[q] = val[base + q];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (r < r1) sort_pairs(C + s, V + s, n);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int64_t q = lane; q < total; q += 64) {
            col[base + q] = C[q];
            val[base + q] = V[q];
        }
    } else if (r < r1) {
        sort_pairs(col + base + s, val + base + s, n);
    }
}

__global__ void k_fill_i32(int32_t* p, int32_t v, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ---- host orchestration --------------------------------------------------

namespace {

struct StaticTables {
    DevBuf<double2> xy;
    DevBuf<double> bbox, part, bpx, bpy, be;
    DevBuf<int32_t> bcell;
    DevBuf<Box> bblk, bsup;
};

int fill_zero(int32_t* p, int64_t n, hipStream_t s)
{
    if (n == 0) return RG_OK;
    k_fill_i32<<<grid_for(n, 256) > 4096 ? 4096 : grid_for(n, 256), 256, 0,
                 s>>>(p, 0, n);
    RG_HIP(hipGetLastError());
    return RG_OK;
}

int prepare_static(const double* x, const double* y, int64_t nx, int64_t ny,
                   StaticTables& T, StaticGrid& G, hipStream_t s)
{
    const int64_t nv = nx * ny;
    const int64_t nbe = 2 * (nx - 1) + 2 * (ny - 1);
    const int64_t nblk = (nbe + BND_BLOCK - 1) / BND_BLOCK;
    const int64_t nsup = (nblk + BND_SUPER - 1) / BND_SUPER;
    int nparts = (int)std::min<int64_t>(grid_for(nv, BBOX_BLOCK), 1024);
    RG_TRY(T.xy.alloc(nv, s));
    RG_TRY(T.bbox.alloc(4, s));
    RG_TRY(T.part.alloc(4 * (size_t)nparts, s));
    RG_TRY(T.bpx.alloc(nbe, s));
    RG_TRY(T.bpy.alloc(nbe, s));
    RG_TRY(T.be.alloc(4 * nbe, s));
    RG_TRY(T.bcell.alloc(nbe, s));
    RG_TRY(T.bblk.alloc(nblk, s));
    RG_TRY(T.bsup.alloc(nsup, s));
    k_interleave<<<std::min<unsigned>(grid_for(nv, 256), 16384), 256, 0, s>>>(
        x, y, nv, T.xy.p);
    k_bbox_partial<<<nparts, BBOX_BLOCK, 0, s>>>(x, y, nv, T.part.p);
    RG_HIP(hipGetLastError());
    k_bbox_final<<<1, BBOX_BLOCK, 0, s>>>(T.part.p, nparts, T.bbox.p);
    RG_HIP(hipGetLastError());
    k_boundary<<<grid_for(nbe, 256), 256, 0, s>>>(x, y, nx, ny, T.bpx.p,
                                                   T.bpy.p, T.be.p, T.bcell.p);
    RG_HIP(hipGetLastError());
    k_edge_boxes<<<grid_for(nblk, 256), 256, 0, s>>>(T.be.p, nbe, T.bblk.p, nblk);
    RG_HIP(hipGetLastError());
    k_super_boxes<<<grid_for(nsup, 256), 256, 0, s>>>(T.bblk.p, nblk, T.bsup.p,
                                                       nsup);
    RG_HIP(hipGetLastError());
    G.x = x;
    G.y = y;
    G.xy = T.xy.p;
    G.nx = nx;
    G.ny = ny;
    G.bbox = T.bbox.p;
    G.bpx = T.bpx.p;
    G.bpy = T.bpy.p;
    G.nbp = nbe;
    G.be = T.be.p;
    G.bcell = T.bcell.p;
    G.nbe = nbe;
    G.bblk = T.bblk.p;
    G.nblk = nblk;
    G.bsup = T.bsup.p;
    G.nsup = nsup;
    return RG_OK;
}

struct PhaseTimer {
    bool on;
    hipStream_t s;
    hipEvent_t ev[10];
    int n = 0;
    explicit PhaseTimer(bool enabled, hipStream_t st) : on(enabled), s(st)
    {
        if (on)
            for (auto& e : ev) (void)hipEventCreate(&e);
    }
    ~PhaseTimer()
    {
        if (on)
            for (auto& e : ev) (void)hipEventDestroy(e);
    }
    void mark()
    {
        if (on && n < 10) (void)hipEventRecord(ev[n++], s);
    }
    double ms(int a, int b)
    {
        if (!on || b >= n) return 0;
        (void)hipEventSynchronize(ev[b]);
        float t = 0;
        (void)hipEventElapsedTime(&t, ev[a], ev[b]);
        return t;
    }
};

}  // namespace

int build_conservative_2d(const double* xi, const double* yi, int64_t nxi,
                          int64_t nyi, const double* xo, const double* yo,
                          int64_t nxo, int64_t nyo, const double* wi,
                          const rg_tuning* tuning, rg_weights** out,
                          hipStream_t s)
{
    *out = nullptr;
    if (injected_failure("build"))
        return fail(RG_ERR_INJECTED, "injected build failure");
    if (!xi || !yi || !xo || !yo)
        return fail(RG_ERR_ARGUMENT, "coordinate pointer is NULL");
    if (nxi < 2 || nyi < 2 || nxo < 2 || nyo < 2)
        return fail(RG_ERR_ARGUMENT,
                    "each grid needs at least 2 vertices along both axes");
    const int64_t n_in = (nxi - 1) * (nyi - 1), n_out = (nxo - 1) * (nyo - 1);
    if (n_in >= INT32_MAX || n_out >= INT32_MAX)
        return fail(RG_ERR_ARGUMENT, "more than 2^31 - 1 cells in one grid");
    int64_t K = 16, stage_per_edge = 8;
    int block = 128;
    bool timings = false;
    if (tuning) {
        if (tuning->edges_per_lane > 0) K = tuning->edges_per_lane;
        if (tuning->block > 0) block = tuning->block;
        if (tuning->stage_per_edge > 0) stage_per_edge = tuning->stage_per_edge;
        timings = (tuning->flags & RG_TUNE_TIMINGS) != 0;
    }
    if (block % 64 != 0 || block > 1024)
        return fail(RG_ERR_ARGUMENT, "block must be a multiple of 64, <= 1024");

    auto t_host0 = std::chrono::steady_clock::now();
    PhaseTimer tm(timings, s);
    tm.mark();  // 0

    // ---- prepare
    DevBuf<double> vol;
    RG_TRY(vol.alloc(n_in, s));
    RG_TRY(launch_area(xi, yi, nxi, nyi, vol.p, s));
    StaticTables tab_in, tab_out;
    StaticGrid st_in, st_out;
    RG_TRY(prepare_static(xi, yi, nxi, nyi, tab_in, st_in, s));
    RG_TRY(prepare_static(xo, yo, nxo, nyo, tab_out, st_out, s));

    Build B;
    B.K = K;
    int64_t chunk = 0, rows = 0, threads = 0;
    for (int p = 0; p < 4; p++) {
        Pass& P = B.pass[p];
        const int sweep_input = p >= 2;
        const int axis = p & 1;
        const double* xs = sweep_input ? xi : xo;
        const double* ys = sweep_input ? yi : yo;
        const int64_t nxs = sweep_input ? nxi : nxo;
        const int64_t nys = sweep_input ? nyi : nyo;
        P.xs = xs;
        P.ys = ys;
        if (axis == 0) {  // aligned view = transpose
            P.sr = 1; P.sk = nys; P.nrows = nys; P.nk = nxs;
        } else {
            P.sr = nys; P.sk = 1; P.nrows = nxs; P.nk = nys;
        }
        P.ncy_sweep = nys - 1;
        P.axis = axis;
        P.sweep_input = sweep_input;
        P.st = sweep_input ? st_out : st_in;
        P.cpr = std::max<int64_t>(1, (P.nk - 1 + K - 1) / K);
        P.chunk_base = chunk;
        P.thread_base = threads;
        B.row_base[p] = rows;
        chunk += P.nrows * P.cpr;
        rows += P.nrows;
        threads += threads_of_pass(P.nrows, P.cpr);
    }
    B.nchunks = chunk;
    B.nrows_total = rows;
    B.nthreads = threads;
    if (B.nthreads >= INT32_MAX)
        return fail(RG_ERR_ARGUMENT, "too many sweep chunks");
    if (B.nthreads % 64 != 0)  // staging regions are whole wavefronts
        return fail(RG_ERR_ARGUMENT, "walk thread count is not wavefront-aligned");
    tm.mark();  // 1

    // ---- start states
    DevBuf<int32_t> start, end, count, hist, counters;
    RG_TRY(start.alloc(B.nchunks, s));
    RG_TRY(end.alloc(B.nchunks, s));
    RG_TRY(count.alloc(B.nchunks, s));
    RG_TRY(hist.alloc(4 * n_in, s));
    // anomalies, fixups, failed/overflow, long runs, restaged chunks
    RG_TRY(counters.alloc(8, s));
    RG_TRY(fill_zero(hist.p, 4 * n_in, s));
    RG_TRY(fill_zero(counters.p, 8, s));
    k_row_start<<<grid_for(B.nrows_total * 64, 256), 256, 0, s>>>(
        B, start.p, counters.p + 0);
    RG_HIP(hipGetLastError());
    k_chunk_start<<<grid_for(B.nthreads, block), block, 0, s>>>(B, start.p);
    RG_HIP(hipGetLastError());
    tm.mark();  // 2

    // ---- count (staging the triples)
    const int64_t cap = K * stage_per_edge;
    DevBuf<Triple> stage;
    RG_TRY(stage.alloc((size_t)B.nthreads * cap, s));
    k_walk_count<<<grid_for(B.nthreads, block), block, 0, s>>>(
        B, start.p, end.p, count.p, hist.p, stage.p, cap);
    RG_HIP(hipGetLastError());
    tm.mark();  // 3

    // ---- fix-up
    k_fixup<<<grid_for(B.nrows_total, 64), 64, 0, s>>>(
        B, start.p, end.p, count.p, hist.p, counters.p + 1, counters.p + 2,
        stage.p, cap);
    RG_HIP(hipGetLastError());
    DevBuf<int64_t> chunk_off, bucket_off;
    RG_TRY(chunk_off.alloc(B.nchunks + 1, s));
    RG_TRY(bucket_off.alloc(4 * n_in + 1, s));
    RG_TRY(scan_i32_to_i64(count.p, chunk_off.p, B.nchunks, s));
    RG_TRY(scan_i32_to_i64(hist.p, bucket_off.p, 4 * n_in, s));
    int64_t tot[2];
    int32_t cnt[4];
    RG_TRY(read_scalars(chunk_off.p + B.nchunks, &tot[0], sizeof(int64_t), s));
    RG_TRY(read_scalars(bucket_off.p + 4 * n_in, &tot[1], sizeof(int64_t), s));
    RG_TRY(read_scalars(counters.p, cnt, sizeof(cnt), s));
    tm.mark();  // 4
    if (cnt[2] != 0)
        return fail(RG_ERR_GEOMETRY,
                    "the sweep did not terminate in " + std::to_string(cnt[2]) +
                        " rows (exactly degenerate crossings)");
    const int64_t T = tot[0];
    if (T != tot[1])
        return fail(RG_ERR_HIP, "internal: triple count " + std::to_string(T) +
                                    " != histogram total " +
                                    std::to_string(tot[1]));
    if (T >= ((int64_t)1 << 32))
        return fail(RG_ERR_ARGUMENT, "more than 2^32 triples in one element");

    // ---- emit
    DevBuf<KV> rec;
    RG_TRY(rec.alloc(T, s));
    uint32_t* cursor = (uint32_t*)hist.p;  // reused: one word per bucket
    k_cursor_init<<<grid_for(4 * n_in, 256), 256, 0, s>>>(bucket_off.p, cursor,
                                                          4 * n_in);
    k_scatter<<<grid_for(B.nthreads, block), block, 0, s>>>(
        B, start.p, count.p, stage.p, cap, chunk_off.p, cursor, rec.p,
        (uint32_t)T, counters.p + 2, counters.p + 4);
    RG_HIP(hipGetLastError());
    k_cursor_check<<<grid_for(4 * n_in, 256), 256, 0, s>>>(bucket_off.p, cursor,
                                                           4 * n_in, counters.p + 2);
    RG_HIP(hipGetLastError());
    stage.reset();
    tm.mark();  // 5

    // ---- reduce -> CSR by input
    DevBuf<int32_t> nnz_row, long_list, group_mode;
    DevBuf<int64_t> row_off;
    const int64_t n_groups = (n_in + 63) / 64;
    RG_TRY(nnz_row.alloc(n_in, s));
    RG_TRY(long_list.alloc(n_in, s));
    RG_TRY(group_mode.alloc(n_groups, s));
    RG_TRY(row_off.alloc(n_in + 1, s));
    k_reduce_groups<<<(unsigned)((n_groups + RED_WAVES - 1) / RED_WAVES),
                      RED_WAVES * 64, 0, s>>>(bucket_off.p, rec.p, vol.p, wi,
                                              n_in, nnz_row.p, group_mode.p,
                                              long_list.p, counters.p + 3);
    k_bucket_long<<<4, 64, 0, s>>>(bucket_off.p, rec.p, long_list.p,
                                   counters.p + 3, nnz_row.p);
    RG_HIP(hipGetLastError());
    RG_TRY(scan_i32_to_i64(nnz_row.p, row_off.p, n_in, s));
    int64_t nnz = 0;
    RG_TRY(read_scalars(row_off.p + n_in, &nnz, sizeof(int64_t), s));
    int32_t cnt2[8];
    RG_TRY(read_scalars(counters.p, cnt2, sizeof(cnt2), s));
    if (cnt2[2] != 0)
        return fail(RG_ERR_HIP, "internal: the emission did not match the "
                                "counted buckets (" + std::to_string(cnt2[2]) +
                                    " mismatches)");
    if (nnz >= INT32_MAX)
        return fail(RG_ERR_ARGUMENT, "more than 2^31 - 1 weights in one element");

    rg_weights* W = new rg_weights();
    W->n_in = n_in;
    W->n_out = n_out;
    W->nnz = nnz;
    W->shape_in[0] = nxi - 1;
    W->shape_in[1] = nyi - 1;
    W->shape_out[0] = nxo - 1;
    W->shape_out[1] = nyo - 1;
    (void)hipGetDevice(&W->device);
    DevBuf<int32_t> rp_in, c_in, rp_out, c_out, hist_out;
    DevBuf<double> v_in, v_out;
    DevBuf<int64_t> off_out;
    int st = RG_OK;
    do {
        if ((st = rp_in.alloc(n_in + 1, s))) break;
        if ((st = c_in.alloc(nnz, s))) break;
        if ((st = v_in.alloc(nnz, s))) break;
        if ((st = rp_out.alloc(n_out + 1, s))) break;
        if ((st = c_out.alloc(nnz, s))) break;
        if ((st = v_out.alloc(nnz, s))) break;
        if ((st = hist_out.alloc(n_out, s))) break;
        if ((st = off_out.alloc(n_out + 1, s))) break;
        if ((st = fill_zero(hist_out.p, n_out, s))) break;
        k_csr_groups<<<grid_for((n_groups + 1) * 64, 256), 256, 0, s>>>(
            bucket_off.p, row_off.p, group_mode.p, rec.p, n_in,
            rp_in.p, c_in.p, v_in.p, hist_out.p);
        if (hipGetLastError() != hipSuccess) {
            st = fail(RG_ERR_HIP, "k_csr_groups launch failed");
            break;
        }
        tm.mark();  // 6
        // ---- transpose -> CSR by output
        if ((st = scan_i32_to_i64(hist_out.p, off_out.p, n_out, s))) break;
        if ((st = fill_zero(hist_out.p, n_out, s))) break;  // cursor
        k_transpose_scatter<<<grid_for(n_in, 128), 128, 0, s>>>(
            rp_in.p, c_in.p, v_in.p, n_in, off_out.p, hist_out.p, c_out.p,
            v_out.p);
        const int64_t sort_groups = n_out / 64 + 1;
        k_rows_sort<<<(unsigned)((sort_groups + SORT_WAVES - 1) / SORT_WAVES),
                      SORT_WAVES * 64, 0, s>>>(off_out.p, n_out, c_out.p,
                                               v_out.p, rp_out.p);
        if (hipGetLastError() != hipSuccess) {
            st = fail(RG_ERR_HIP, "transpose launch failed");
            break;
        }
        tm.mark();  // 7
    } while (0);
    if (st != RG_OK) {
        delete W;
        return st;
    }
    W->rowptr_in = rp_in.release();
    W->col_in = c_in.release();
    W->val_in = v_in.release();
    W->rowptr_out = rp_out.release();
    W->col_out = c_out.release();
    W->val_out = v_out.release();
    rg_build_stats& S = W->stats;
    S.n_triples = T;
    S.n_chunks = B.nchunks;
    S.n_fixups = cnt[1];
    S.n_search_anomalies = cnt[0];
    S.n_restaged = cnt2[4];
    // every inside sub-segment emits two triples except on the first and
    // last row of a sweep, which emit one
    S.n_subsegments = T / 2;
    if (timings) {
        S.ms_prepare = tm.ms(0, 1);
        S.ms_start = tm.ms(1, 2);
        S.ms_count = tm.ms(2, 3);
        S.ms_fixup = tm.ms(3, 4);
        S.ms_emit = tm.ms(4, 5);
        S.ms_reduce = tm.ms(5, 6);
        S.ms_transpose = tm.ms(6, 7);
        S.ms_total = tm.ms(0, 7);
    }
    (void)t_host0;
    *out = W;
    return RG_OK;
}

void weights_release(rg_weights* W)
{
    if (!W) return;
    dev_free(W->rowptr_in);
    dev_free(W->col_in);
    dev_free(W->val_in);
    dev_free(W->rowptr_out);
    dev_free(W->col_out);
    dev_free(W->val_out);
    dev_free(W->val_in_f32);
    dev_free(W->val_out_f32);
    for (auto& pl : W->plan) {
        dev_free(pl.box);
        dev_free(pl.off);
    }
    delete W;
}

// ---- export / import -----------------------------------------------------

__global__ void k_export(const int32_t* __restrict__ rowptr,
                         const int32_t* __restrict__ col,
                         const double* __restrict__ v, int64_t nrows,
                         int64_t* __restrict__ oi, int64_t* __restrict__ oo,
                         double* __restrict__ ov)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= nrows) return;
    for (int32_t q = rowptr[i]; q < rowptr[i + 1]; q++) {
        oi[q] = i;
        oo[q] = col[q];
        ov[q] = v[q];
    }
}

// Import: group the storage-ordered triples by one index, keeping storage
// order inside each group (positions are the sort keys).
__global__ void k_hist_i64(const int64_t* __restrict__ idx, int64_t n,
                           int32_t* __restrict__ hist, int32_t* __restrict__ bad,
                           int64_t limit)
{
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * blockDim.x) {
        int64_t k = idx[q];
        if (k < 0 || k >= limit) {
            atomicAdd(bad, 1);
            continue;
        }
        atomicAdd(&hist[k], 1);
    }
}

__global__ void k_check_range(const int64_t* __restrict__ idx, int64_t n,
                              int32_t* __restrict__ bad, int64_t limit)
{
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * blockDim.x) {
        int64_t k = idx[q];
        if (k < 0 || k >= limit) atomicAdd(bad, 1);
    }
}

__global__ void k_group_scatter(const int64_t* __restrict__ idx, int64_t n,
                                const int64_t* __restrict__ off,
                                int32_t* __restrict__ cursor,
                                int32_t* __restrict__ pos_out)
{
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * blockDim.x) {
        int64_t k = idx[q];
        int64_t p = off[k] + atomicAdd(&cursor[k], 1);
        if (p < off[k + 1]) pos_out[p] = (int32_t)q;
    }
}

__global__ void k_group_finish(const int64_t* __restrict__ off, int64_t nrows,
                               int32_t* __restrict__ pos,
                               const int64_t* __restrict__ other,
                               const double* __restrict__ w,
                               int32_t* __restrict__ rowptr,
                               int32_t* __restrict__ col,
                               double* __restrict__ val)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i > nrows) return;
    rowptr[i] = (int32_t)off[i];
    if (i == nrows) return;
    const int64_t s = off[i], n = off[i + 1] - s;
    // order the group by storage position (insertion for short groups)
    int32_t* p = pos + s;
    if (n <= 48) {
        for (int64_t a = 1; a < n; a++) {
            int32_t k = p[a];
            int64_t b = a - 1;
            while (b >= 0 && p[b] > k) {
                p[b + 1] = p[b];
                b--;
            }
            p[b + 1] = k;
        }
    } else {
        double* dummy = val + s;  // reuse as the carried payload
        for (int64_t a = 0; a < n; a++) dummy[a] = 0;
        sort_pairs(p, dummy, n);
    }
    for (int64_t a = 0; a < n; a++) {
        int32_t q = p[a];
        col[s + a] = (int32_t)other[q];
        val[s + a] = w[q];
    }
}

namespace {

int group_csr(const int64_t* idx, const int64_t* other, const double* w,
              int64_t nnz, int64_t nrows, int64_t ncols, int32_t** rowptr,
              int32_t** col, double** val, hipStream_t s)
{
    DevBuf<int32_t> hist, pos, bad, rp, c;
    DevBuf<int64_t> off;
    DevBuf<double> v;
    RG_TRY(hist.alloc(nrows, s));
    RG_TRY(bad.alloc(2, s));
    RG_TRY(off.alloc(nrows + 1, s));
    RG_TRY(pos.alloc(nnz, s));
    RG_TRY(rp.alloc(nrows + 1, s));
    RG_TRY(c.alloc(nnz, s));
    RG_TRY(v.alloc(nnz, s));
    RG_TRY(fill_zero(hist.p, nrows, s));
    RG_TRY(fill_zero(bad.p, 2, s));
    unsigned gs = std::min<unsigned>(grid_for(nnz, 256), 8192);
    k_hist_i64<<<gs, 256, 0, s>>>(idx, nnz, hist.p, bad.p, nrows);
    k_check_range<<<gs, 256, 0, s>>>(other, nnz, bad.p + 1, ncols);
    RG_HIP(hipGetLastError());
    int32_t nbad[2];
    RG_TRY(read_scalars(bad.p, nbad, sizeof(nbad), s));
    if (nbad[0] || nbad[1])
        return fail(RG_ERR_ARGUMENT, "weight index out of range for the given "
                                     "cell counts");
    RG_TRY(scan_i32_to_i64(hist.p, off.p, nrows, s));
    RG_TRY(fill_zero(hist.p, nrows, s));
    k_group_scatter<<<gs, 256, 0, s>>>(idx, nnz, off.p, hist.p, pos.p);
    k_group_finish<<<grid_for(nrows + 1, 128), 128, 0, s>>>(
        off.p, nrows, pos.p, other, w, rp.p, c.p, v.p);
    RG_HIP(hipGetLastError());
    *rowptr = rp.release();
    *col = c.release();
    *val = v.release();
    return RG_OK;
}

}  // namespace

int import_triples(const int64_t* ii, const int64_t* io, const double* w,
                   int64_t nnz, int64_t n_in, int64_t n_out, rg_weights** out,
                   hipStream_t s)
{
    *out = nullptr;
    if (nnz < 0 || n_in < 0 || n_out < 0)
        return fail(RG_ERR_ARGUMENT, "negative size");
    if (nnz >= INT32_MAX || n_in >= INT32_MAX || n_out >= INT32_MAX)
        return fail(RG_ERR_ARGUMENT, "sizes beyond 2^31 - 1");
    if (nnz > 0 && (!ii || !io || !w))
        return fail(RG_ERR_ARGUMENT, "weight array is NULL");
    rg_weights* W = new rg_weights();
    W->n_in = n_in;
    W->n_out = n_out;
    W->nnz = nnz;
    W->shape_in[1] = n_in;
    W->shape_out[1] = n_out;
    (void)hipGetDevice(&W->device);
    int st = group_csr(io, ii, w, nnz, n_out, n_in, &W->rowptr_out, &W->col_out,
                       &W->val_out, s);
    if (st == RG_OK)
        st = group_csr(ii, io, w, nnz, n_in, n_out, &W->rowptr_in, &W->col_in,
                       &W->val_in, s);
    if (st != RG_OK) {
        weights_release(W);
        return st;
    }
    *out = W;
    return RG_OK;
}

int export_triples(const rg_weights* W, int64_t* oi, int64_t* oo, double* ov,
                   hipStream_t s)
{
    if (W->nnz == 0) return RG_OK;
    k_export<<<grid_for(W->n_in, 128), 128, 0, s>>>(W->rowptr_in, W->col_in,
                                                     W->val_in, W->n_in, oi, oo,
                                                     ov);
    RG_HIP(hipGetLastError());
    return RG_OK;
}

}  // namespace rg

using namespace rg;

extern "C" {

int rg_build_conservative_2d(const double* xi, const double* yi, int64_t nxi,
                             int64_t nyi, const double* xo, const double* yo,
                             int64_t nxo, int64_t nyo, const double* wi,
                             const rg_tuning* tuning, rg_weights_t* out,
                             void* stream)
{
    if (!out) return fail(RG_ERR_ARGUMENT, "out is NULL");
    return build_conservative_2d(xi, yi, nxi, nyi, xo, yo, nxo, nyo, wi, tuning,
                                 out, (hipStream_t)stream);
}

int rg_weights_info(rg_weights_t w, int64_t* n_in, int64_t* n_out,
                    int64_t* nnz)
{
    if (!w) return fail(RG_ERR_ARGUMENT, "weights handle is NULL");
    if (n_in) *n_in = w->n_in;
    if (n_out) *n_out = w->n_out;
    if (nnz) *nnz = w->nnz;
    return RG_OK;
}

int rg_weights_stats(rg_weights_t w, rg_build_stats* stats)
{
    if (!w || !stats) return fail(RG_ERR_ARGUMENT, "NULL argument");
    *stats = w->stats;
    return RG_OK;
}

int rg_weights_export(rg_weights_t w, int64_t* indices_input,
                      int64_t* indices_output, double* values, void* stream)
{
    if (!w) return fail(RG_ERR_ARGUMENT, "weights handle is NULL");
    if (w->nnz && (!indices_input || !indices_output || !values))
        return fail(RG_ERR_ARGUMENT, "output array is NULL");
    return export_triples(w, indices_input, indices_output, values,
                          (hipStream_t)stream);
}

int rg_weights_import(const int64_t* indices_input,
                      const int64_t* indices_output, const double* values,
                      int64_t nnz, int64_t n_in, int64_t n_out,
                      rg_weights_t* out, void* stream)
{
    if (!out) return fail(RG_ERR_ARGUMENT, "out is NULL");
    return import_triples(indices_input, indices_output, values, nnz, n_in,
                          n_out, out, (hipStream_t)stream);
}

int rg_weights_set_shape(rg_weights_t w, int64_t ncx_in, int64_t ncy_in,
                         int64_t ncx_out, int64_t ncy_out)
{
    if (!w) return fail(RG_ERR_ARGUMENT, "weights handle is NULL");
    if (ncx_in * ncy_in != w->n_in || ncx_out * ncy_out != w->n_out)
        return fail(RG_ERR_ARGUMENT, "shapes do not match the cell counts");
    w->shape_in[0] = ncx_in;
    w->shape_in[1] = ncy_in;
    w->shape_out[0] = ncx_out;
    w->shape_out[1] = ncy_out;
    return RG_OK;
}

int rg_weights_free(rg_weights_t w)
{
    weights_release(w);
    return RG_OK;
}

}  // extern "C"
