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
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "rg_internal.hpp"
#include "rg_sweep.hpp"
#include "rg_wave.hpp"

namespace rg {

// ---- atomics usable from the shared walk code ----------------------------

// Triples are bucketed by sweep pass and by group of 64 consecutive input
// cells (every pair's cell space is padded to a multiple of 64, so a group
// belongs to one pair), pass-major: bucket (pass p, group g) at slot
// p * G + g (G groups in all).  The histogram of buckets is small enough to
// stay in cache, a walk's consecutive triples mostly fall in one bucket, and
// the scatter claims whole runs of records per bucket.  Each record carries
// its cell within the group in a side array of bytes; the reduction splits
// the group by cell.

// Histogram increments are combined in registers: a walk's consecutive
// triples mostly fall in one or two (cell, pass) buckets (the static cell of
// a sweep over the output grid; the left and right cells of an input edge),
// so two pending (bucket, count) entries replace most atomics.
struct HistCache {
    int32_t b0 = -1, b1 = -1;  // bucket indices (a build has fewer than 2^31 buckets)
    int32_t n0 = 0, n1 = 0;
    __device__ void add(int32_t* __restrict__ hist, int32_t b, int32_t d)
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
// flush() after the walk.  The staged triples are stored nontemporal: the
// staging (0.8 GB at 2049^2 -> 1537^2) is read once, by the scatter, long
// after, and written through the caches it evicts the grid lines the walks
// of neighbouring chunks re-read (the walk 0.82 -> 0.79 ms there).
struct StageSink {
    int32_t n;
    int64_t g, cap;
    Triple* __restrict__ stage;
    int32_t* __restrict__ hist;
    int32_t p;
    int64_t ps;  // bucket (pass p, cell c) is hist[p * ps + (c >> 6)]
    HistCache hc;
    __device__ void emit(int64_t cin, int64_t cout, double a)
    {
        if (n < cap) {
            Triple t;
            t.cin = (int32_t)cin;
            t.cout = (int32_t)cout;
            t.a = a;
            typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
            __builtin_nontemporal_store(__builtin_bit_cast(u32x4, t),
                                        (u32x4*)&stage[stage_index(g, n, cap)]);
        }
        n++;
        hc.add(hist, (int32_t)(p * ps + (cin >> 6)), 1);
    }
    __device__ void flush() { hc.flush(hist); }
};

struct UncountSink {
    int32_t* hist;
    int32_t p;
    int64_t ps;
    HistCache hc;
    __device__ void emit(int64_t cin, int64_t, double)
    {
        hc.add(hist, (int32_t)(p * ps + (cin >> 6)), -1);
    }
    __device__ void flush() { hc.flush(hist); }
};

// Bucket cursors: cursor[b] starts at the first record of bucket b and each
// store claims records with a returning atomic (record positions fit in 32
// bits: a build holds fewer than 2^32 triples).  A store is only
// bounds-checked against the record count; k_cursor_check then verifies that
// every cursor ended exactly at its bucket's end, which holds iff each
// bucket received exactly the triples the count walk histogrammed.
__device__ inline void put_record(KV* __restrict__ rec, uint8_t* __restrict__ rcell,
                                  uint32_t pos, int64_t cin, int64_t cout, uint32_t ordinal,
                                  double a)
{
    KV r;
    r.key = ((uint64_t)(uint32_t)cout << 32) | (uint64_t)ordinal;
    r.val = a;
    rec[pos] = r;
    rcell[pos] = (uint8_t)(cin & 63);
}

// Keeps the triples beyond the first `cap` of a chunk (the ones the
// counting walk could not stage) in the chunk's range of the overflow
// buffer: its first triple is the chunk's triple number `cap`.
struct OverflowSink {
    int32_t n;
    int64_t cap;
    Triple* __restrict__ out;
    int64_t room;  // triples the chunk's range can take
    __device__ void emit(int64_t cin, int64_t cout, double a)
    {
        if (n >= cap && n - cap < room) {
            Triple t;
            t.cin = (int32_t)cin;
            t.cout = (int32_t)cout;
            t.a = a;
            out[n - cap] = t;
        }
        n++;
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
constexpr int64_t VTILE = 4096;  // vertices per prepare tile

// A static grid of a (batched) build and where its derived tables start in
// the concatenated arrays.
struct GridDesc {
    const double* x;
    const double* y;
    int64_t nx, ny;
    int64_t v_base;     // first vertex in xy_all
    int64_t e_base;     // first boundary edge (bpx, bpy, be, bcell)
    int64_t blk_base;   // first edge block box
    int64_t sup_base;   // first super-block box
    int64_t tile_base;  // first vertex tile
    int64_t pyr_base;   // first polygon-edge run (y range)
};

// The cells of one pair of a batched build.
struct PairCells {
    const double* xi;
    const double* yi;
    const double2* xyi;  // the input grid's interleaved vertex copy
    int64_t nyi;
    const double* wi;
    int64_t cell_gbase, n_in;  // input cells: first (padded) global id, count
    int64_t out_gbase, n_out;  // output cells
};

// unit -> segment lookup table (pair of a wavefront, row, cell group ...)
__global__ void k_segments(const int64_t* __restrict__ prefix, int n,
                           int64_t nunits, int32_t* __restrict__ out)
{
    const int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (u < nunits) out[u] = seg_of(prefix, n, u);
}

// One block per tile of VTILE vertices of one grid: interleaves the
// vertices into (x, y) pairs and reduces the tile's bounding box.
__global__ __launch_bounds__(BBOX_BLOCK) void k_grid_tiles(
    const GridDesc* __restrict__ grids, const int64_t* __restrict__ tile_prefix,
    int ngrids, double2* __restrict__ xy_all, double* __restrict__ part)
{
    const int64_t tile = blockIdx.x;
    const GridDesc& G = grids[seg_of(tile_prefix, ngrids, tile)];
    const int64_t nv = G.nx * G.ny;
    const int64_t v0 = (tile - G.tile_base) * VTILE;
    const int64_t v1 = v0 + VTILE < nv ? v0 + VTILE : nv;
    double v[4] = {G.x[v0], G.y[v0], G.x[v0], G.y[v0]};
    for (int64_t i = v0 + threadIdx.x; i < v1; i += BBOX_BLOCK) {
        const double a = G.x[i], c = G.y[i];
        double2 q;
        q.x = a;
        q.y = c;
        xy_all[G.v_base + i] = q;
        v[0] = nan_min(v[0], a);
        v[1] = nan_min(v[1], c);
        v[2] = nan_max(v[2], a);
        v[3] = nan_max(v[3], c);
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
        for (int q = 0; q < 4; q++) part[4 * tile + q] = sh[q][0];
}

// One block per grid: the bounding box from its tiles' partial boxes.
__global__ __launch_bounds__(BBOX_BLOCK) void k_grid_bbox(
    const int64_t* __restrict__ tile_prefix, const double* __restrict__ part,
    double* __restrict__ bbox_all)
{
    const int g = blockIdx.x;
    const int64_t t0 = tile_prefix[g], t1 = tile_prefix[g + 1];
    double v[4] = {part[4 * t0], part[4 * t0 + 1], part[4 * t0 + 2], part[4 * t0 + 3]};
    for (int64_t t = t0 + threadIdx.x; t < t1; t += BBOX_BLOCK) {
        v[0] = nan_min(v[0], part[4 * t + 0]);
        v[1] = nan_min(v[1], part[4 * t + 1]);
        v[2] = nan_max(v[2], part[4 * t + 2]);
        v[3] = nan_max(v[3], part[4 * t + 3]);
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
        for (int q = 0; q < 4; q++) bbox_all[4 * g + q] = sh[q][0];
}

// Boundary polygon and boundary edge table of every grid (each grid has
// 2(nx-1)+2(ny-1) of both).
__global__ void k_grid_boundary(const GridDesc* __restrict__ grids,
                                const int64_t* __restrict__ e_prefix, int ngrids,
                                int64_t ne_all, double* __restrict__ bpx,
                                double* __restrict__ bpy, double* __restrict__ be,
                                int32_t* __restrict__ bcell)
{
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= ne_all) return;
    const GridDesc& G = grids[seg_of(e_prefix, ngrids, e)];
    const int64_t le = e - G.e_base;
    const int64_t v = boundary_vertex(G.nx, G.ny, le);
    bpx[e] = G.x[v];
    bpy[e] = G.y[v];
    boundary_edge(G.x, G.y, G.nx, G.ny, le, be + 4 * e, bcell + e);
}

// The y range of the vertices of each run of 64 boundary-polygon edges
// (edge k runs from vertex k - 1 to vertex k, edge 0 from the last vertex):
// a run whose range excludes a point's y adds nothing to its winding number.
__global__ void k_grid_poly_yranges(const GridDesc* __restrict__ grids,
                                    const int64_t* __restrict__ pyr_prefix, int ngrids,
                                    int64_t npyr_all, const double* __restrict__ bpy_all,
                                    double2* __restrict__ pyr)
{
    // one wavefront per run, one lane per edge
    const int64_t b = grid_wave();
    const int lane = threadIdx.x & 63;
    if (b >= npyr_all) return;  // uniform per wavefront
    const GridDesc& G = grids[seg_of(pyr_prefix, ngrids, b)];
    const int64_t nbp = 2 * (G.nx - 1) + 2 * (G.ny - 1);
    const int64_t k = (b - G.pyr_base) * 64 + lane;
    const double* bpy = bpy_all + G.e_base;
    double lo = __builtin_huge_val(), hi = -__builtin_huge_val();
    if (k < nbp) {
        const int64_t km = k == 0 ? nbp - 1 : k - 1;
        // fmin/fmax ignore NaN; an edge with a NaN end adds nothing
        lo = fmin(bpy[km], bpy[k]);
        hi = fmax(bpy[km], bpy[k]);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        lo = fmin(lo, __shfl_xor(lo, o, 64));
        hi = fmax(hi, __shfl_xor(hi, o, 64));
    }
    if (lane == 0) pyr[b] = make_double2(lo, hi);
}

__global__ void k_grid_edge_boxes(const GridDesc* __restrict__ grids,
                                  const int64_t* __restrict__ blk_prefix,
                                  int ngrids, int64_t nblk_all,
                                  const double* __restrict__ be_all,
                                  Box* __restrict__ bblk)
{
    const int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (b >= nblk_all) return;
    const GridDesc& G = grids[seg_of(blk_prefix, ngrids, b)];
    const int64_t nbe = 2 * (G.nx - 1) + 2 * (G.ny - 1);
    const int64_t lb = b - G.blk_base;
    const double* be = be_all + 4 * G.e_base;
    Box r;
    r.x0 = r.y0 = __builtin_huge_val();
    r.x1 = r.y1 = -__builtin_huge_val();
    const int64_t e1 = (lb + 1) * BND_BLOCK < nbe ? (lb + 1) * BND_BLOCK : nbe;
    for (int64_t e = lb * BND_BLOCK; e < e1; e++) {
        const double* q = be + 4 * e;
        // fmin/fmax ignore NaN; an edge with a NaN end never hits anyway
        r.x0 = fmin(r.x0, fmin(q[0], q[2]));
        r.x1 = fmax(r.x1, fmax(q[0], q[2]));
        r.y0 = fmin(r.y0, fmin(q[1], q[3]));
        r.y1 = fmax(r.y1, fmax(q[1], q[3]));
    }
    bblk[b] = r;
}

__global__ void k_grid_super_boxes(const GridDesc* __restrict__ grids,
                                   const int64_t* __restrict__ sup_prefix,
                                   int ngrids, int64_t nsup_all,
                                   const Box* __restrict__ bblk_all,
                                   Box* __restrict__ bsup)
{
    const int64_t sidx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (sidx >= nsup_all) return;
    const GridDesc& G = grids[seg_of(sup_prefix, ngrids, sidx)];
    const int64_t nbe = 2 * (G.nx - 1) + 2 * (G.ny - 1);
    const int64_t nblk = (nbe + BND_BLOCK - 1) / BND_BLOCK;
    const int64_t ls = sidx - G.sup_base;
    const Box* bblk = bblk_all + G.blk_base;
    Box r;
    r.x0 = r.y0 = __builtin_huge_val();
    r.x1 = r.y1 = -__builtin_huge_val();
    const int64_t b1 = (ls + 1) * BND_SUPER < nblk ? (ls + 1) * BND_SUPER : nblk;
    for (int64_t b = ls * BND_SUPER; b < b1; b++) {
        r.x0 = fmin(r.x0, bblk[b].x0);
        r.y0 = fmin(r.y0, bblk[b].y0);
        r.x1 = fmax(r.x1, bblk[b].x1);
        r.y1 = fmax(r.y1, bblk[b].y1);
    }
    bsup[sidx] = r;
}

// One thread per grid: the box of all its boundary edges (NaN ends ignored,
// as in the block boxes) -- a sweep edge that misses it misses every block.
__global__ void k_grid_boundary_box(const int64_t* __restrict__ sup_prefix, int ngrids,
                                    const Box* __restrict__ bsup, Box* __restrict__ ball)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngrids) return;
    Box r;
    r.x0 = r.y0 = __builtin_huge_val();
    r.x1 = r.y1 = -__builtin_huge_val();
    for (int64_t q = sup_prefix[g]; q < sup_prefix[g + 1]; q++) {
        r.x0 = fmin(r.x0, bsup[q].x0);
        r.y0 = fmin(r.y0, bsup[q].y0);
        r.x1 = fmax(r.x1, bsup[q].x1);
        r.y1 = fmax(r.y1, bsup[q].y1);
    }
    ball[g] = r;
}

// Input cell areas (and the input weighting, 1 where a pair has none) over
// the padded global cell space; padding cells get area 1 and weight 1 (they
// receive no triples).
__global__ void k_pair_cells(const PairCells* __restrict__ pcs,
                             const int32_t* __restrict__ pair_of_group,
                             int64_t ncells_all, double* __restrict__ vol,
                             double* __restrict__ wi_all)
{
    const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (c >= ncells_all) return;
    const PairCells& P = pcs[pair_of_group[c >> 6]];
    const int64_t lc = c - P.cell_gbase;
    double a = 1, w = 1;
    if (lc < P.n_in) {
        const int64_t ncy = P.nyi - 1;
        a = cell_area_xy(P.xyi, P.nyi, lc / ncy, lc % ncy);
        if (P.wi) w = P.wi[lc];
    }
    vol[c] = a;
    if (wi_all) wi_all[c] = w;
}

// ---- start states --------------------------------------------------------

// Row start of the reference (:293-322) for rows_per_wave rows per
// wavefront (lane j < rows_per_wave owns row j): bounding box; the winding
// number over the static boundary, one row at a time split across the
// lanes; the secant search, every owned row at once on its own lane; the
// exhaustive fallback, one row at a time spread over the wave.  The host
// puts one row per wavefront while the rows alone do not fill the GPU, and
// more when they do: a wavefront per row ran each secant search on one lane
// (the 256-pair batch's build, about a million rows: 20.24 -> 20.04 ms).
__global__ void k_row_start(const Build* __restrict__ builds,
                            const int32_t* __restrict__ pair_of_row,
                            int64_t nrows_all, int32_t* __restrict__ start,
                            int32_t* __restrict__ anomalies, int rows_per_wave)
{
    const int R = rows_per_wave;  // 1..64, a power of two
    const int64_t r0 = grid_wave() * R;
    const int lane = threadIdx.x & 63;
    if (r0 >= nrows_all) return;  // uniform per wavefront
    // lane j < R owns row r0 + j
    const int64_t rr = r0 + lane;
    const bool own = lane < R && rr < nrows_all;
    int32_t k_own = 0, p_own = 0;
    int64_t row_own = 0;
    double px_own = 0, py_own = 0;
    bool box_own = false;
    if (own) {
        k_own = pair_of_row[rr];
        const Build& B = builds[k_own];
        const int64_t lr = rr - B.row_gbase;
        p_own = pass_of_row(B, lr);
        const Pass& P = B.pass[p_own];
        row_own = lr - B.row_base[p_own];
        px_own = sweep_x(P, row_own, 0);
        py_own = sweep_y(P, row_own, 0);
        const double* bb = P.st.bbox;
        box_own = bb[0] <= px_own && px_own <= bb[2] && bb[1] <= py_own && py_own <= bb[3];
    }
    // the winding number of each owned row, over its static boundary split
    // across the lanes: a sum of +-1 and +-0.5 terms, exact in any order;
    // runs of 64 edges whose vertices all lie strictly above or below py
    // add nothing (their y - py never changes sign or is zero)
    double w_own = 0;
    for (int j = 0; j < R; j++) {  // uniform
        if (!__builtin_amdgcn_readlane((int)box_own, j)) continue;
        const Build& B = builds[__builtin_amdgcn_readlane(k_own, j)];
        const StaticGrid& st = B.pass[__builtin_amdgcn_readlane(p_own, j)].st;
        const double px = __shfl(px_own, j, 64), py = __shfl(py_own, j, 64);
        double w = 0;
        for (int64_t b0 = 0; b0 < st.npyr; b0 += 64) {  // uniform
            const int64_t b = b0 + lane;
            bool touch = false;
            if (b < st.npyr) {
                const double2 r = st.pyr[b];
                touch = !(py < r.x || py > r.y);
            }
            uint64_t m = __ballot(touch);
            while (m) {
                const int64_t k = (b0 + __builtin_ctzll(m)) * 64 + lane;
                m &= m - 1;
                if (k < st.nbp) {
                    const int64_t km = k == 0 ? st.nbp - 1 : k - 1;
                    w += winding_term(px, py, st.bpx[km], st.bpy[km], st.bpx[k], st.bpy[k]);
                }
            }
        }
        w = wave_sum(w);
        if (lane == j) w_own = w;
    }
    // the owned rows inside their boundary: the secant search, one lane per
    // row, all at once
    int64_t c_own = -1;
    if (own && box_own && w_own != 0) {
        const StaticGrid& st = builds[k_own].pass[p_own].st;
        c_own = index_secant_xy(st.xy, st.nx, st.ny, px_own, py_own);
    }
    // the rows the secant search did not settle: the exhaustive search, one
    // row at a time spread over the wave
    uint64_t mb = __ballot(own && box_own && w_own != 0 && c_own == RG_SEARCH_BRUTE);
    while (mb) {
        const int j = __builtin_ctzll(mb);
        mb &= mb - 1;
        const Build& B = builds[__builtin_amdgcn_readlane(k_own, j)];
        const StaticGrid& st = B.pass[__builtin_amdgcn_readlane(p_own, j)].st;
        const int64_t c = wave_brute(st.x, st.y, st.nx, st.ny, __shfl(px_own, j, 64),
                                     __shfl(py_own, j, 64), lane);
        if (lane == j) c_own = c;
    }
    if (own) {
        const Build& B = builds[k_own];
        const Pass& P = B.pass[p_own];
        int32_t state = STATE_OUTSIDE;
        if (box_own && w_own != 0) {
            if (c_own >= 0)
                state = (int32_t)c_own;
            else  // inside the boundary polygon but in no cell: the reference
                  // then indexes with sys.maxsize; treated as outside here
                atomicAdd(anomalies, 1);
        }
        start[B.chunk_gbase + P.chunk_base + row_own * P.cpr] = state;
    }
}

// The walk kernels take blocks of up to 512 threads; the bound (rather than
// the default 1024) lets the walk keep its state in registers without
// spilling.
constexpr int WALK_MAX_BLOCK = 512;

// The pair of walk thread t of a batched build (wavefront-uniform: every
// pair's walk threads are whole wavefronts).
__device__ inline const Build& build_of_thread(const Build* __restrict__ builds,
                                               const int32_t* __restrict__ pair_of_wave,
                                               int64_t t)
{
    return builds[__builtin_amdgcn_readfirstlane(pair_of_wave[t >> 6])];
}

// The pass of a walk thread's chunk, read as a wave-uniform value: each pass
// (and each pair) starts its threads at a multiple of 64, so the active lanes
// of a wavefront share one pass, and the pass table's fields then go through
// the scalar cache into scalar registers.
__device__ inline const Pass& pass_of_wave(const Build& B, const ChunkRef& r)
{
    return B.pass[__builtin_amdgcn_readfirstlane(r.p)];
}

__global__ __launch_bounds__(WALK_MAX_BLOCK) void k_chunk_start(const Build* __restrict__ builds,
                              const int32_t* __restrict__ pair_of_wave,
                              int64_t nthreads_all, int32_t* __restrict__ start,
                              int guess_outside)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nthreads_all) return;
    const Build& B = build_of_thread(builds, pair_of_wave, t);
    const ChunkRef r = chunk_of_thread(B, t - B.thread_gbase);
    if (!r.valid || r.c == 0) return;
    start[B.chunk_gbase + r.g] =
        guess_outside ? STATE_OUTSIDE : chunk_start_guess(pass_of_wave(B, r), r.row, r.c * B.K);
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

__global__ __launch_bounds__(WALK_MAX_BLOCK) void k_walk_count(const Build* __restrict__ builds,
                             const int32_t* __restrict__ pair_of_wave,
                             int64_t nthreads_all, int64_t ngroups_all,
                             const int32_t* __restrict__ start,
                             int32_t* __restrict__ end,
                             int32_t* __restrict__ count,
                             int32_t* __restrict__ hist,
                             Triple* __restrict__ stage, int64_t cap)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nthreads_all) return;
    const Build& B = build_of_thread(builds, pair_of_wave, t);
    const ChunkRef r = chunk_of_thread(B, t - B.thread_gbase);
    if (!r.valid) return;
    const Pass& P = pass_of_wave(B, r);
    const int64_t g = B.chunk_gbase + r.g;
    int64_t k0, k1;
    chunk_range(B, P, r.c, &k0, &k1);
    StageSink sink{0, t, cap, stage, hist + (B.cell_gbase >> 6), r.p, ngroups_all, {}};
    end[g] = walk_chunk(B, P, r.row, k0, k1, start[g], sink);
    sink.flush();
    count[g] = sink.n;
}

// One wavefront per row: make every chunk start where its predecessor
// ended.  The lanes check the row's chunk links with coalesced loads; only a
// row with a mismatched link or a failed chunk is re-walked, by lane 0, in
// chunk order (a fix can change the end state the next link compares with).
__global__ __launch_bounds__(WALK_MAX_BLOCK) void k_fixup(const Build* __restrict__ builds,
                        const int32_t* __restrict__ pair_of_row,
                        int64_t nrows_all, int64_t ngroups_all,
                        int32_t* __restrict__ start,
                        int32_t* __restrict__ end, int32_t* __restrict__ count,
                        int32_t* __restrict__ hist_all, int32_t* __restrict__ nfix,
                        int32_t* __restrict__ failed,
                        Triple* __restrict__ stage, int64_t cap)
{
    const int64_t gr = grid_wave();
    const int lane = threadIdx.x & 63;
    if (gr >= nrows_all) return;  // uniform per wavefront
    const Build& B = builds[pair_of_row[gr]];
    const int64_t r = gr - B.row_gbase;
    const int p = pass_of_row(B, r);
    const Pass& P = B.pass[p];
    const int64_t row = r - B.row_base[p];
    const int64_t g0 = B.chunk_gbase + P.chunk_base + row * P.cpr;
    bool bad = false;
    for (int64_t c = lane; c < P.cpr; c += 64) {
        if (end[g0 + c] == STATE_FAILED) bad = true;
        if (c >= 1 && start[g0 + c] != end[g0 + c - 1]) bad = true;
    }
    if (__ballot(bad) == 0 || lane != 0) return;
    int32_t* hist = hist_all + (B.cell_gbase >> 6);
    int32_t fixes = 0;
    for (int64_t c = 1; c < P.cpr; c++) {
        const int64_t g = g0 + c;
        const int32_t truth = end[g - 1];
        if (truth == STATE_FAILED) break;
        if (start[g] == truth) continue;
        int64_t k0, k1;
        chunk_range(B, P, c, &k0, &k1);
        UncountSink undo{hist, p, ngroups_all, {}};
        (void)walk_chunk(B, P, row, k0, k1, start[g], undo);
        undo.flush();
        StageSink redo{0, B.thread_gbase + thread_of_chunk(P, row, c), cap, stage,
                       hist, p, ngroups_all, {}};
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

// Triples beyond the staging capacity, per chunk (then scanned into the
// chunks' ranges of the overflow buffer).
__global__ void k_overflow_counts(const int32_t* __restrict__ count, int64_t nchunks,
                                  int64_t cap, int32_t* __restrict__ ov,
                                  unsigned long long* __restrict__ total)
{
    const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int64_t v = 0;
    if (g < nchunks) {
        v = count[g] > cap ? count[g] - cap : 0;
        ov[g] = (int32_t)v;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(total, (unsigned long long)v);
}

// Walks every chunk that overflowed the staging again and keeps the
// triples the staging could not hold.
__global__ __launch_bounds__(WALK_MAX_BLOCK) void k_restage(const Build* __restrict__ builds,
                          const int32_t* __restrict__ pair_of_wave,
                          int64_t nthreads_all,
                          const int32_t* __restrict__ start,
                          const int32_t* __restrict__ count,
                          const int64_t* __restrict__ ov_off,
                          Triple* __restrict__ overflow, int64_t cap, int64_t ov_cap,
                          int32_t* __restrict__ restaged)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nthreads_all) return;
    const Build& B = build_of_thread(builds, pair_of_wave, t);
    const ChunkRef r = chunk_of_thread(B, t - B.thread_gbase);
    if (!r.valid) return;
    const int64_t g = B.chunk_gbase + r.g;
    if (count[g] <= cap) return;
    const Pass& P = pass_of_wave(B, r);
    int64_t k0, k1;
    chunk_range(B, P, r.c, &k0, &k1);
    // (a capacity-mode build sizes the buffer in advance: a chunk past its
    // end keeps nothing, and the build's flag reports it)
    const int64_t room = ov_cap - ov_off[g];
    OverflowSink sink{0, cap, overflow + ov_off[g], room > 0 ? room : 0};
    (void)walk_chunk(B, P, r.row, k0, k1, start[g], sink);
    atomicAdd(restaged, 1);
}

// Orders the LDS accesses of a wavefront's lanes before and after it.
__device__ inline void wave_sync_lds()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Exclusive prefix over the wave of one value per lane.
__device__ inline int32_t wave_exclusive(int32_t v, int lane)
{
    int32_t incl = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int32_t y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
    }
    return incl - v;
}

// Staged triples -> buckets, each tagged with its emission ordinal (the
// chunk's first ordinal plus its position in the chunk) and its cell byte.
// A wavefront takes its lanes' triples SCAT_R per lane per round and groups
// them by bucket in an LDS hash table (one entry per bucket met in the
// round: bucket, count, claimed base, offset in the round's staging); one
// lane per entry claims the bucket's records with a single atomic.  The
// round's records are then laid out in LDS bucket by bucket and written
// out by consecutive lanes to consecutive addresses, so every bucket's run
// goes to memory as whole lines rather than one 16-byte store per lane at
// scattered addresses.  A full table falls back to one claim and one store
// per triple.  A chunk's triples beyond the staging capacity come from the
// overflow buffer (k_restage).
constexpr int SCAT_R = 8;     // triples per lane per round
constexpr int SCAT_HT = 256;  // hash entries per wavefront
constexpr int SCAT_N = 64 * SCAT_R;
constexpr int SCAT_BLOCK = 128;  // 2 wavefronts, 27 KiB of LDS per block

struct ScatTable {
    int32_t key[SCAT_HT];
    int32_t cnt[SCAT_HT];
    uint32_t base[SCAT_HT];
    int32_t off[SCAT_HT];
    uint64_t rk[SCAT_N];  // the round's records, bucket by bucket
    double rv[SCAT_N];
    uint8_t rslot[SCAT_N];  // their table entry
    uint8_t rc[SCAT_N];     // their cell byte
};

__global__ __launch_bounds__(WALK_MAX_BLOCK) void k_scatter(const Build* __restrict__ builds,
                          const int32_t* __restrict__ pair_of_wave,
                          int64_t nthreads_all, int64_t ngroups_all,
                          const int32_t* __restrict__ count,
                          const Triple* __restrict__ stage, int64_t cap,
                          const Triple* __restrict__ ovbuf,
                          const int64_t* __restrict__ ov_off, int64_t ov_cap,
                          const int64_t* __restrict__ chunk_off,
                          uint32_t* __restrict__ cursor_all,
                          KV* __restrict__ rec, uint8_t* __restrict__ rcell,
                          uint32_t n_rec,
                          int32_t* __restrict__ overflow)
{
    extern __shared__ unsigned char scat_lds[];
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nthreads_all) return;  // nthreads_all is a multiple of 64: whole waves
    const int lane = threadIdx.x & 63;
    ScatTable& H = ((ScatTable*)scat_lds)[threadIdx.x >> 6];
    const Build& B = build_of_thread(builds, pair_of_wave, t);
    const ChunkRef r = chunk_of_thread(B, t - B.thread_gbase);
    const int64_t g = B.chunk_gbase + r.g;
    // every lane of the wave sweeps one pass: the bucket row is wave-uniform
    uint32_t* cursor = cursor_all + (B.cell_gbase >> 6) + (int64_t)r.p * ngroups_all;
    int32_t n = r.valid ? count[g] : 0;
    const uint32_t ord = r.valid ? (uint32_t)chunk_off[g] : 0;
    const Triple* ov = n > cap ? ovbuf + ov_off[g] : nullptr;
    if (n > cap) {  // only what the overflow buffer holds (all of it unless capped)
        const int64_t room = ov_cap - ov_off[g];
        if ((int64_t)n - cap > room) n = (int32_t)(cap + (room > 0 ? room : 0));
    }
    int32_t mx = n;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = max(mx, __shfl_xor(mx, o, 64));
    for (int h = lane; h < SCAT_HT; h += 64) {
        H.key[h] = -1;
        H.cnt[h] = 0;
    }
    wave_sync_lds();
    // the next round's triples are requested while this round is placed
    Triple en[SCAT_R];
#pragma unroll
    for (int k = 0; k < SCAT_R; k++)
        if (k < n) en[k] = k < cap ? stage[stage_index(t, k, cap)] : ov[k - cap];
    for (int32_t q0 = 0; q0 < mx; q0 += SCAT_R) {
        Triple e[SCAT_R];
#pragma unroll
        for (int k = 0; k < SCAT_R; k++) e[k] = en[k];
#pragma unroll
        for (int k = 0; k < SCAT_R; k++) {
            const int32_t q = q0 + SCAT_R + k;
            if (q < n) en[k] = q < cap ? stage[stage_index(t, q, cap)] : ov[q - cap];
        }
        int32_t slot[SCAT_R], local[SCAT_R];
#pragma unroll
        for (int k = 0; k < SCAT_R; k++) {
            slot[k] = -2;  // no triple
            if (q0 + k < n) {
                const int32_t b = e[k].cin >> 6;
                int32_t h = b & (SCAT_HT - 1);
                slot[k] = -1;  // table full: claim alone
                for (int probe = 0; probe < SCAT_HT; probe++) {
                    const int32_t cur = atomicCAS(&H.key[h], -1, b);
                    if (cur == -1 || cur == b) {
                        slot[k] = h;
                        break;
                    }
                    h = (h + 1) & (SCAT_HT - 1);
                }
                if (slot[k] >= 0) local[k] = atomicAdd(&H.cnt[h], 1);
            }
        }
        wave_sync_lds();
        // claims, and each entry's offset in the round's staging (a wave
        // prefix over the table, SCAT_HT / 64 entries per lane); the claims'
        // results are kept in registers and written to the table only after
        // the round's records are laid out, so their latency overlaps the
        // layout
        int32_t c4[SCAT_HT / 64], sum = 0;
        uint32_t b4[SCAT_HT / 64];
#pragma unroll
        for (int u = 0; u < SCAT_HT / 64; u++) {
            const int h = lane * (SCAT_HT / 64) + u;
            c4[u] = H.key[h] >= 0 ? H.cnt[h] : 0;
            b4[u] = 0;
            if (c4[u] > 0) b4[u] = atomicAdd(&cursor[H.key[h]], (uint32_t)c4[u]);
            sum += c4[u];
        }
        int32_t o = wave_exclusive(sum, lane);
        const int32_t staged = __shfl(o + sum, 63, 64);
#pragma unroll
        for (int u = 0; u < SCAT_HT / 64; u++) {
            H.off[lane * (SCAT_HT / 64) + u] = o;
            o += c4[u];
        }
        wave_sync_lds();
#pragma unroll
        for (int k = 0; k < SCAT_R; k++) {
            if (slot[k] == -2) continue;
            const uint64_t key = ((uint64_t)(uint32_t)e[k].cout << 32) | (uint64_t)(ord + (uint32_t)(q0 + k));
            if (slot[k] >= 0) {
                const int32_t i = H.off[slot[k]] + local[k];
                H.rk[i] = key;
                H.rv[i] = e[k].a;
                H.rslot[i] = (uint8_t)slot[k];
                H.rc[i] = (uint8_t)(e[k].cin & 63);
            } else {
                const uint32_t pos = atomicAdd(&cursor[e[k].cin >> 6], 1u);
                if (pos < n_rec)
                    put_record(rec, rcell, pos, e[k].cin, e[k].cout, ord + (uint32_t)(q0 + k),
                               e[k].a);
                else
                    atomicAdd(overflow, 1);
            }
        }
#pragma unroll
        for (int u = 0; u < SCAT_HT / 64; u++)
            if (c4[u] > 0) H.base[lane * (SCAT_HT / 64) + u] = b4[u];
        wave_sync_lds();
        for (int32_t i = lane; i < staged; i += 64) {
            const int h = H.rslot[i];
            const uint32_t pos = H.base[h] + (uint32_t)(i - H.off[h]);
            if (pos < n_rec) {
                KV rr;
                rr.key = H.rk[i];
                rr.val = H.rv[i];
                rec[pos] = rr;
                rcell[pos] = H.rc[i];
            } else {
                atomicAdd(overflow, 1);
            }
        }
        wave_sync_lds();
        for (int h = lane; h < SCAT_HT; h += 64) {
            H.key[h] = -1;
            H.cnt[h] = 0;
        }
        wave_sync_lds();
    }
}

// ---- reduce and CSR ------------------------------------------------------

// Coalescing, one wavefront per group of 64 consecutive input cells.  The
// group's records sit in four pass ranges (the buckets are pass-major), in
// claim order, each with its cell byte.  The wave counts them per cell in
// LDS (the cell-major offsets of the group's output, written to cell_off),
// places them cell-major in LDS, and each lane divides its cell's terms by
// the cell area, sorts them and reduces equal outputs (bucket_reduce); the
// wave writes the group's unique (out, weight) pairs packed, in row order,
// from the group's first offset.  A group beyond the LDS share is first
// copied cell-major into `red` (global), then goes through LDS in runs of
// consecutive cells, each cell reduced in place at its own offset (the
// group is then marked unpacked); a single cell beyond the share is left to
// k_bucket_long.
// Bits of the flag word (counters[8]) of a capacity-mode build (graph
// capture): which fixed capacity the data exceeded, or that the build failed.
constexpr int32_t CAP_TRIPLES = 1, CAP_OVERFLOW = 2, CAP_WEIGHTS = 4, CAP_FAILED = 8,
                  CAP_MISMATCH = 16;

// wavefronts per block: one, so that the blocks pack the LDS more finely (a
// block's 16.6 KiB per wavefront: nine wavefronts per CU instead of eight
// in blocks of two; 0.547 -> 0.511 ms at 2049^2 -> 1537^2, the 256-pair
// batch's reduce phase 7.31 -> 7.11 ms)
constexpr int RED_WAVES = 1;
constexpr int HEAVY_WAVES = 2;  // k_reduce_heavy's
constexpr int RED_CAP = 1024;    // LDS entries per wavefront (16 KiB)
constexpr int RED_NET = 16;      // largest cell sorted in registers
// the lane-strided layout of a group whose cells all fit RED_NET entries
// puts entry j of cell c at j * 64 + c: RED_NET rows of 64 must fit
static_assert(RED_CAP >= RED_NET * 64, "the lane-strided layout needs RED_NET * 64 entries");

// One cell's terms: divided by the cell area (and weighted), sorted by key
// in registers through a network of N, written back in place.
template <int N, int ST = 1>
__device__ inline void sort_cell_in_registers(uint64_t* K, double* V, int32_t n, double vol,
                                              const double* wi, int64_t i)
{
    uint64_t k[N];
    double v[N];
#pragma unroll
    for (int j = 0; j < N; j++) {
        k[j] = ~(uint64_t)0;  // padding sorts last
        v[j] = 0;
        if (j < n) {
            k[j] = K[j * ST];
            v[j] = weight_of(V[j * ST], vol, wi, i);
        }
    }
    sort_network<N, uint64_t>(k, v);
#pragma unroll
    for (int j = 0; j < N; j++)
        if (j < n) {
            K[j * ST] = k[j];
            V[j * ST] = v[j];
        }
}

// Records of one group (all passes): the group totals that give the
// group-major output offsets.
__global__ void k_group_counts(const int64_t* __restrict__ bucket_off, int64_t n_groups,
                               int32_t* __restrict__ cnt)
{
    const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (g >= n_groups) return;
    int64_t t = 0;
#pragma unroll
    for (int p = 0; p < 4; p++)
        t += bucket_off[p * n_groups + g + 1] - bucket_off[p * n_groups + g];
    cnt[g] = (int32_t)t;
}

// One reduced entry (output cell `out`, weight) of a group: stored with its
// arrival rank among the entries of its output row (the high word of the
// key), counted in hist_out for the transposed CSR.
__device__ inline void put_reduced(KV* __restrict__ dst, uint64_t out, double w,
                                   int32_t* __restrict__ hist_out)
{
    const uint32_t rank = (uint32_t)atomicAdd(&hist_out[(int64_t)out], 1);
    KV r;
    r.key = out | ((uint64_t)rank << 32);
    r.val = w;
    *dst = r;
}

// put_reduced for the m reduced entries of one cell (keys = output cells),
// PUT_R rank atomics in flight at a time.
constexpr int PUT_R = 8;

__device__ inline void put_reduced_cell(KV* __restrict__ dst, const uint64_t* K,
                                        const double* V, int64_t m,
                                        int32_t* __restrict__ hist_out, int64_t st = 1)
{
    for (int64_t t0 = 0; t0 < m; t0 += PUT_R) {
        uint32_t rk[PUT_R];
#pragma unroll
        for (int u = 0; u < PUT_R; u++)
            if (t0 + u < m) rk[u] = (uint32_t)atomicAdd(&hist_out[(int64_t)K[(t0 + u) * st]], 1);
#pragma unroll
        for (int u = 0; u < PUT_R; u++)
            if (t0 + u < m) {
                KV r;
                r.key = K[(t0 + u) * st] | ((uint64_t)rk[u] << 32);
                r.val = V[(t0 + u) * st];
                dst[t0 + u] = r;
            }
    }
}

// The records of one group, reached through its four pass ranges: their
// total (the group's count, of which group_off is the exclusive scan) and
// the end of the last range, which ends last (the records sit pass-major).
struct GroupRecs {
    int64_t ps0, ps1, ps2, ps3, e0, e1, e2, total, end;
    __device__ GroupRecs(const int64_t* __restrict__ bucket_off, int64_t n_groups, int64_t grp)
    {
        ps0 = bucket_off[grp];
        ps1 = bucket_off[n_groups + grp];
        ps2 = bucket_off[2 * n_groups + grp];
        ps3 = bucket_off[3 * n_groups + grp];
        e0 = bucket_off[grp + 1] - ps0;
        e1 = e0 + bucket_off[n_groups + grp + 1] - ps1;
        e2 = e1 + bucket_off[2 * n_groups + grp + 1] - ps2;
        end = bucket_off[3 * n_groups + grp + 1];
        total = e2 + end - ps3;
    }
    // position of record q of the concatenated pass ranges
    __device__ int64_t src(int64_t q) const
    {
        return q < e0 ? ps0 + q : q < e1 ? ps1 + (q - e0) : q < e2 ? ps2 + (q - e1)
                                                               : ps3 + (q - e2);
    }
};

// Groups whose records fit the LDS share; a larger group is listed for
// k_reduce_heavy.
__global__ __launch_bounds__(RED_WAVES * 64) void k_reduce_groups(
    const int64_t* __restrict__ bucket_off, int64_t n_groups,
    const int64_t* __restrict__ group_off, const KV* __restrict__ rec,
    const uint8_t* __restrict__ rcell, KV* __restrict__ red,
    int64_t* __restrict__ cell_off, const double* __restrict__ vol_in,
    const double* __restrict__ wi, int32_t* __restrict__ nnz_row,
    int32_t* __restrict__ group_mode, int32_t* __restrict__ long_list,
    int32_t* __restrict__ n_long, int32_t* __restrict__ heavy_list,
    int32_t* __restrict__ n_heavy, const PairCells* __restrict__ pcs,
    const int32_t* __restrict__ pair_of_group, int32_t* __restrict__ hist_out_all,
    int64_t red_cap, int32_t* __restrict__ cap_flag)
{
    __shared__ uint64_t sk[RED_WAVES][RED_CAP];
    __shared__ double sv[RED_WAVES][RED_CAP];
    __shared__ int32_t scnt[RED_WAVES][64];
    const int wid = block_wave(), lane = threadIdx.x & 63;
    if (blockIdx.x == 0 && threadIdx.x == 0) cell_off[n_groups * 64] = group_off[n_groups];
    const int64_t grp = (int64_t)blockIdx.x * RED_WAVES + wid;
    if (grp >= n_groups) return;  // uniform per wavefront; no block barriers below
    // requested first, with the group's offsets: the cell areas and the
    // pair's output-cell base, which the records do not decide
    const double vol = vol_in[grp * 64 + lane];
    int32_t* const hist_out =
        hist_out_all + pcs[__builtin_amdgcn_readfirstlane(pair_of_group[grp])].out_gbase;
    // every offset of the group takes part in the first test, so they are
    // requested together (a load first used after a test is not issued
    // before it: one round trip per test)
    const int64_t base = group_off[grp];
    const GroupRecs G(bucket_off, n_groups, grp);
    const int64_t total = G.total;
    if (base + total > red_cap || G.end > red_cap) {
        // past the records a capacity-mode build holds: nothing of it is kept
        nnz_row[grp * 64 + lane] = 0;
        cell_off[grp * 64 + lane] = 0;
        if (lane == 0) {
            group_mode[grp] = 0;
            atomicOr(cap_flag, CAP_TRIPLES);
        }
        return;
    }
    if (total > RED_CAP) {
        if (lane == 0) heavy_list[atomicAdd(n_heavy, 1)] = (int32_t)grp;
        return;
    }
    const int64_t i = grp * 64 + lane;
    int32_t* C = scnt[wid];
    // the records are read once, into registers (record q = lane + 64 u of
    // the concatenated pass ranges), with nontemporal loads: nothing reads
    // them again, and through the caches they would evict the cell bytes
    // and offsets of the groups in flight (0.518 -> 0.503 ms at 2049^2 ->
    // 1537^2)
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    KV rr[RED_CAP / 64];
    int32_t cc[RED_CAP / 64];
    C[lane] = 0;
    wave_sync_lds();
#pragma unroll
    for (int u = 0; u < RED_CAP / 64; u++) {
        const int64_t q = lane + 64 * u;
        cc[u] = -1;
        if (q < total) {
            const int64_t src = G.src(q);
            cc[u] = rcell[src];
            rr[u] = __builtin_bit_cast(KV, __builtin_nontemporal_load((const u32x4*)&rec[src]));
        }
    }
    // (the count's returning atomic also gives each record its place in its
    // cell, used by both layouts below)
    int32_t dd[RED_CAP / 64];
#pragma unroll
    for (int u = 0; u < RED_CAP / 64; u++) dd[u] = cc[u] >= 0 ? atomicAdd(&C[cc[u]], 1) : 0;
    wave_sync_lds();
    // records per cell, and the cell-major offsets
    const int32_t n = C[lane];
    const int32_t s = wave_exclusive(n, lane);
    cell_off[i] = base + s;
    uint64_t* K = sk[wid];
    double* V = sv[wid];
    int32_t nmax = n;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) nmax = max(nmax, __shfl_xor(nmax, o, 64));
    if (nmax <= RED_NET) {
        // every cell fits RED_NET entries: entry j of cell c sits at
        // j * 64 + c, so the lanes (one per cell) touch consecutive LDS
        // words at every step of the sort and the reduction
#pragma unroll
        for (int u = 0; u < RED_CAP / 64; u++)
            if (cc[u] >= 0) {
                K[dd[u] * 64 + cc[u]] = rr[u].key;
                V[dd[u] * 64 + cc[u]] = rr[u].val;
            }
        wave_sync_lds();
        if (nmax <= RED_NET / 2)
            sort_cell_in_registers<RED_NET / 2, 64>(K + lane, V + lane, n, vol, wi, i);
        else
            sort_cell_in_registers<RED_NET, 64>(K + lane, V + lane, n, vol, wi, i);
        bool lr1 = false;  // (a run of at most RED_NET terms is never long)
        const int64_t m1 = bucket_reduce_sorted(K + lane, V + lane, n, &lr1, 64);
        const int64_t dst1 = wave_exclusive((int32_t)m1, lane);
        put_reduced_cell(red + base + dst1, K + lane, V + lane, m1, hist_out, 64);
        nnz_row[i] = (int32_t)m1;
        if (lane == 0) group_mode[grp] = 0;
        return;
    }
    wave_sync_lds();
    // the keys cell-major in LDS; then every record's place in its cell is
    // the number of smaller keys of its cell (the keys are unique: they hold
    // the emission ordinal), and the records are written there, weighted
#pragma unroll
    for (int u = 0; u < RED_CAP / 64; u++) {
        const int32_t cs = __shfl(s, cc[u] >= 0 ? cc[u] : lane, 64);
        if (cc[u] >= 0) K[cs + dd[u]] = rr[u].key;
    }
    wave_sync_lds();
    int32_t pos[RED_CAP / 64];
#pragma unroll
    for (int u = 0; u < RED_CAP / 64; u++) {
        const int32_t ce = cc[u] >= 0 ? cc[u] : lane;
        const int32_t cs = __shfl(s, ce, 64), cn = __shfl(n, ce, 64);
        pos[u] = -1;
        if (cc[u] >= 0) {
            const uint64_t k = rr[u].key;
            // four keys requested before any is compared
            int32_t r = 0, j = cs;
            const int32_t je = cs + cn;
            for (; j + 4 <= je; j += 4) {
                const uint64_t k0 = K[j], k1 = K[j + 1], k2 = K[j + 2], k3 = K[j + 3];
                r += (k0 < k ? 1 : 0) + (k1 < k ? 1 : 0) + (k2 < k ? 1 : 0) + (k3 < k ? 1 : 0);
            }
            for (; j < je; j++) r += K[j] < k ? 1 : 0;
            pos[u] = cs + r;
        }
    }
    wave_sync_lds();
#pragma unroll
    for (int u = 0; u < RED_CAP / 64; u++) {
        const int32_t ce = cc[u] >= 0 ? cc[u] : lane;
        const double cv = __shfl(vol, ce, 64);
        if (pos[u] >= 0) {
            K[pos[u]] = rr[u].key;
            V[pos[u]] = weight_of(rr[u].val, cv, wi, grp * 64 + ce);
        }
    }
    wave_sync_lds();
    bool lr = false;
    int64_t m = bucket_reduce_sorted(K + s, V + s, n, &lr);
    if (lr) m = n;  // left sorted, uncompacted, for k_bucket_long
    const bool packed = __ballot(lr) == 0;
    // packed destination: exclusive prefix of m over the wave; a group
    // with a long run keeps the cell starts
    const int64_t dst = packed ? wave_exclusive((int32_t)m, lane) : s;
    if (lr) {
        for (int64_t t = 0; t < m; t++) {  // finished by k_bucket_long
            KV r;
            r.key = K[s + t];
            r.val = V[s + t];
            red[base + dst + t] = r;
        }
    } else {
        put_reduced_cell(red + base + dst, K + s, V + s, m, hist_out);
    }
    nnz_row[i] = (int32_t)m;
    if (lr) long_list[atomicAdd(n_long, 1)] = (int32_t)i;
    if (lane == 0) group_mode[grp] = packed ? 0 : 1;
}

// Groups beyond the LDS share, one wavefront each.  The waves are resident
// and pull the next group of the list from a counter (a work queue: the
// heavy groups differ in size by orders of magnitude, so a fixed
// assignment leaves waves idle behind the one that drew the largest).  The
// cells' counts give their cell-major offsets in `red` (the group is marked
// unpacked); the cells are then taken in runs of consecutive cells that fit
// the LDS share.  For each run the wave places the records of the run's
// cells in LDS, cell-major; every record finds its place in its cell's key
// order by counting the smaller keys of its cell (the keys are unique: they
// hold the emission ordinal), and each lane reduces one cell at its own
// offset.  A single cell beyond the share is copied cell-major into `red`
// and left to k_bucket_long.
//
// Each run lists its records' sources in LDS first (the group's cell bytes
// read HEAVY_UNROLL at a time per lane), then loads the records with
// HEAVY_UNROLL loads per lane in flight, each record written over its own
// source slot: two short dependent chains per run instead of one per
// record.
constexpr int HEAVY_CAP = RED_CAP;  // LDS entries per wavefront
constexpr int HEAVY_UNROLL = 4;  // records loaded per lane at once
constexpr int HEAVY_SCAN = 8;    // cell bytes requested per lane at once
// A group's cell bytes are kept in LDS from the counting scan when it has at
// most this many records (most heavy groups), so the scans of its runs read
// LDS instead of memory; 2 KiB per wavefront keeps four blocks per CU.
constexpr int HEAVY_CELLB = 2048;

__global__ __launch_bounds__(HEAVY_WAVES * 64) void k_reduce_heavy(
    const int64_t* __restrict__ bucket_off, int64_t n_groups,
    const int64_t* __restrict__ group_off, const KV* __restrict__ rec,
    const uint8_t* __restrict__ rcell, KV* __restrict__ red,
    int64_t* __restrict__ cell_off, const double* __restrict__ vol_in,
    const double* __restrict__ wi, int32_t* __restrict__ nnz_row,
    int32_t* __restrict__ group_mode, int32_t* __restrict__ long_list,
    int32_t* __restrict__ n_long, const int32_t* __restrict__ heavy_list,
    const int32_t* __restrict__ n_heavy, int32_t* __restrict__ heavy_next,
    const PairCells* __restrict__ pcs, const int32_t* __restrict__ pair_of_group,
    int32_t* __restrict__ hist_out_all)
{
    __shared__ uint64_t sk[HEAVY_WAVES][HEAVY_CAP];
    __shared__ double sv[HEAVY_WAVES][HEAVY_CAP];
    __shared__ uint8_t sc[HEAVY_WAVES][HEAVY_CAP];
    __shared__ uint8_t scb[HEAVY_WAVES][HEAVY_CELLB];
    __shared__ int32_t scnt[HEAVY_WAVES][64];
    const int wid = block_wave(), lane = threadIdx.x & 63;
    int32_t* C = scnt[wid];
    uint64_t* K = sk[wid];
    double* V = sv[wid];
    int64_t* SRC = reinterpret_cast<int64_t*>(sv[wid]);  // a run's record sources
    uint8_t* SC = sc[wid];
    uint8_t* CB = scb[wid];  // the group's cell bytes, by record
    constexpr int U = HEAVY_CAP / 64;
    const int32_t nh = *n_heavy;
    // the next list position, requested one group ahead
    int32_t hn = 0;
    if (lane == 0) hn = atomicAdd(heavy_next, 1);
    for (;;) {
        // (all lanes are active here: the first active lane is lane 0, and
        // the group's tables are then read with scalar loads)
        const int32_t h = __builtin_amdgcn_readfirstlane(hn);
        if (h >= nh) break;  // uniform per wavefront
        if (lane == 0) hn = atomicAdd(heavy_next, 1);
        const int64_t grp = heavy_list[h];
        const int64_t i = grp * 64 + lane;
        const double vol = vol_in[i];
        const int64_t base = group_off[grp];
        const GroupRecs G(bucket_off, n_groups, grp);
        const int64_t total = G.total;
        int32_t* hist_out =
            hist_out_all + pcs[__builtin_amdgcn_readfirstlane(pair_of_group[grp])].out_gbase;
        const bool cached = total <= HEAVY_CELLB;
        C[lane] = 0;
        wave_sync_lds();
        for (int64_t q0 = lane; q0 < total; q0 += 64 * HEAVY_SCAN) {
            int32_t cb[HEAVY_SCAN];
#pragma unroll
            for (int u = 0; u < HEAVY_SCAN; u++)
                cb[u] = q0 + 64 * u < total ? rcell[G.src(q0 + 64 * u)] : -1;
#pragma unroll
            for (int u = 0; u < HEAVY_SCAN; u++)
                if (cb[u] >= 0) {
                    atomicAdd(&C[cb[u]], 1);
                    if (cached) CB[q0 + 64 * u] = (uint8_t)cb[u];
                }
        }
        wave_sync_lds();
        const int32_t n = C[lane];
        const int32_t s = wave_exclusive(n, lane);
        cell_off[i] = base + s;
        wave_sync_lds();
        int c = 0;  // first cell of the next run
        while (c < 64) {
            const int32_t sb = __shfl(s, c, 64);  // wave-uniform
            const bool in = lane >= c && s + n - sb <= HEAVY_CAP;
            const uint64_t mask = __ballot(in);
            if (mask == 0) {
                // a cell beyond the LDS share: its records go cell-major to
                // `red`, finished by k_bucket_long (flagged in the long list)
                if (lane == 0) C[0] = 0;
                wave_sync_lds();
                for (int64_t q = lane; q < total; q += 64) {
                    const int64_t src = G.src(q);
                    if ((cached ? CB[q] : rcell[src]) == c)
                        red[base + sb + atomicAdd(&C[0], 1)] = rec[src];
                }
                if (lane == c) {  // (k_bucket_long takes the cell's size from nnz_row)
                    nnz_row[i] = n;
                    long_list[atomicAdd(n_long, 1)] = (int32_t)i | INT32_MIN;
                }
                wave_sync_lds();
                c++;
                continue;
            }
            const int d = c + __popcll(mask);  // cells [c, d) fit
            const int32_t cnt = __shfl(s + n, d - 1, 64) - sb;
            C[lane] = s - sb;  // placement cursors of the run's cells
            wave_sync_lds();
            // the run's records, cell-major in LDS: their sources first ...
            for (int64_t q0 = lane; q0 < total; q0 += 64 * HEAVY_SCAN) {
                int32_t cb[HEAVY_SCAN];
#pragma unroll
                for (int u = 0; u < HEAVY_SCAN; u++) {
                    const int64_t q = q0 + 64 * u;
                    cb[u] = q < total ? (cached ? CB[q] : rcell[G.src(q)]) : -1;
                }
#pragma unroll
                for (int u = 0; u < HEAVY_SCAN; u++)
                    if (cb[u] >= c && cb[u] < d) {
                        const int32_t p = atomicAdd(&C[cb[u]], 1);
                        SRC[p] = G.src(q0 + 64 * u);
                        SC[p] = (uint8_t)cb[u];
                    }
            }
            wave_sync_lds();
            // ... then the records, each written over its own source slot
            for (int32_t e0 = lane; e0 < cnt; e0 += 64 * HEAVY_UNROLL) {
                KV r[HEAVY_UNROLL];
#pragma unroll
                for (int u = 0; u < HEAVY_UNROLL; u++) {
                    const int32_t e = e0 + 64 * u;
                    if (e < cnt) r[u] = rec[SRC[e]];
                }
#pragma unroll
                for (int u = 0; u < HEAVY_UNROLL; u++) {
                    const int32_t e = e0 + 64 * u;
                    if (e < cnt) {
                        K[e] = r[u].key;
                        V[e] = r[u].val;
                    }
                }
            }
            wave_sync_lds();
            // each record's place in its cell: the number of smaller keys
            uint64_t ke[U];
            double ve[U];
            int32_t pe[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int32_t e = lane + 64 * u;
                const int32_t ce = e < cnt ? SC[e] : lane;
                const int32_t cs = __shfl(s, ce, 64) - sb;
                const int32_t cn = __shfl(n, ce, 64);
                const double cv = __shfl(vol, ce, 64);
                pe[u] = -1;
                if (e < cnt) {
                    const uint64_t k = K[e];
                    // four keys requested before any is compared (one LDS
                    // latency per four keys, not per key)
                    int32_t r = 0, j = cs;
                    const int32_t je = cs + cn;
                    for (; j + 4 <= je; j += 4) {
                        const uint64_t k0 = K[j], k1 = K[j + 1], k2 = K[j + 2], k3 = K[j + 3];
                        r += (k0 < k ? 1 : 0) + (k1 < k ? 1 : 0) + (k2 < k ? 1 : 0) + (k3 < k ? 1 : 0);
                    }
                    for (; j < je; j++) r += K[j] < k ? 1 : 0;
                    ke[u] = k;
                    ve[u] = weight_of(V[e], cv, wi, grp * 64 + ce);
                    pe[u] = cs + r;
                }
            }
            wave_sync_lds();
#pragma unroll
            for (int u = 0; u < U; u++)
                if (pe[u] >= 0) {
                    K[pe[u]] = ke[u];
                    V[pe[u]] = ve[u];
                }
            wave_sync_lds();
            // the run's cells reduced, each written packed after the run's
            // previous cells (cell_off moves to it): the CSR pass then reads
            // the run's entries as one dense range instead of one short
            // range per cell's records
            uint64_t* Kc = K + (s - sb);
            double* Vc = V + (s - sb);
            bool lrc = false;
            int64_t m = 0;
            if (in) {
                m = bucket_reduce_sorted(Kc, Vc, n, &lrc);
                if (lrc) m = n;  // left sorted, uncompacted, for k_bucket_long
            }
            const int64_t dst = base + sb + wave_exclusive(in ? (int32_t)m : 0, lane);
            if (in) {
                if (lrc) {
                    for (int64_t t = 0; t < m; t++) {
                        KV r;
                        r.key = Kc[t];
                        r.val = Vc[t];
                        red[dst + t] = r;
                    }
                } else {
                    put_reduced_cell(red + dst, Kc, Vc, m, hist_out);
                }
                cell_off[i] = dst;
                nnz_row[i] = (int32_t)m;
                if (lrc) long_list[atomicAdd(n_long, 1)] = (int32_t)i;
            }
            wave_sync_lds();
            c = d;
        }
        if (lane == 0) group_mode[grp] = 1;
        wave_sync_lds();  // LDS reuse by the next group
    }
}

// Cells the group reduction left over: a cell beyond the LDS share (flag
// bit set; its records already sit cell-major in `red`) is weighted,
// sorted and reduced here; a cell with a run longer than the LDS reduction
// takes (left sorted and weighted in `red`) is reduced here.
__global__ void k_bucket_long(const int64_t* __restrict__ cell_off, KV* __restrict__ red,
                              const double* __restrict__ vol_in,
                              const double* __restrict__ wi,
                              const int32_t* __restrict__ long_list,
                              const int32_t* __restrict__ n_long,
                              int32_t* __restrict__ nnz_row,
                              const PairCells* __restrict__ pcs,
                              const int32_t* __restrict__ pair_of_group,
                              int32_t* __restrict__ hist_out_all)
{
    const int32_t nl = *n_long;
    for (int32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < nl;
         q += gridDim.x * blockDim.x) {
        const int32_t e = long_list[q];
        const int64_t i = e & INT32_MAX;
        // (a long cell keeps all its records: nnz_row holds its record count
        // until the cell is reduced here)
        const int64_t start = cell_off[i], n = nnz_row[i];
        KV* r = red + start;
        if (e < 0) {
            const double vol = vol_in[i];
            for (int64_t t = 0; t < n; t++) r[t].val = weight_of(r[t].val, vol, wi, i);
            sort_records(r, n);
        }
        const int64_t m = bucket_reduce_long(r, n);
        nnz_row[i] = (int32_t)m;
        int32_t* hist_out = hist_out_all + pcs[pair_of_group[i >> 6]].out_gbase;
        for (int64_t t = 0; t < m; t++) put_reduced(r + t, r[t].key, r[t].val, hist_out);
    }
}

// Both CSRs from the reduced groups.  By input: the wave copies the group's
// entries in order, each read from its row's offset (one contiguous range
// for a packed group).  By output: each
// entry goes to its output row at the row's offset plus its arrival rank
// (counted by the reductions); k_rows_sort then orders each row by column,
// which makes the result deterministic.  Also writes each pair's int32 row
// pointers by input (relative to the pair's first entry; pair k's n_in + 1
// pointers start at rowptr_all[cell_gbase + k]).
__global__ void k_csr_groups(const PairCells* __restrict__ pcs,
                             const int32_t* __restrict__ pair_of_group,
                             int64_t n_groups,
                             const int64_t* __restrict__ cell_off,
                             const int64_t* __restrict__ row_off,
                             const int32_t* __restrict__ group_mode,
                             const KV* __restrict__ rec,
                             int32_t* __restrict__ rowptr_all,
                             int32_t* __restrict__ col,
                             double* __restrict__ out_val,
                             const int64_t* __restrict__ t_off,
                             int32_t* __restrict__ t_col,
                             double* __restrict__ t_val, int64_t nnz_cap)
{
    const int64_t grp = grid_wave();
    const int lane = threadIdx.x & 63;
    if (grp >= n_groups) return;  // uniform per wavefront
    const int k = __builtin_amdgcn_readfirstlane(pair_of_group[grp]);
    const PairCells& P = pcs[k];
    const int64_t i0 = grp * 64, iend = P.cell_gbase + P.n_in;
    const int64_t i1 = i0 + 64 < iend ? i0 + 64 : iend;
    // a capacity-mode build past its weight capacity keeps nothing more
    // (its flag reports it); an ordinary build passes its own weight count
    if (row_off[i1 > i0 ? i1 : i0] > nnz_cap) return;
    const int64_t ebase = row_off[P.cell_gbase];
    int32_t* rowptr = rowptr_all + P.cell_gbase + k;
    const int64_t i = i0 + lane;
    if (i < i1) rowptr[i - P.cell_gbase] = (int32_t)(row_off[i] - ebase);
    if (i == iend - 1) rowptr[P.n_in] = (int32_t)(row_off[iend] - ebase);
    if (i0 >= i1) return;
    const int64_t* toff = t_off + P.out_gbase;
    const int64_t d0 = row_off[i0];
    const int64_t cnt = row_off[i1] - d0;
    // the row of entry q: the lanes' rows bracket the group's entries
    const int64_t my_d = i < i1 ? row_off[i] - d0 : INT64_MAX;
    // where the row's entries sit in `rec`, less the row's first entry: the
    // group's start when packed, the row's own start when unpacked
    const int64_t my_src = group_mode[grp] == 0 ? cell_off[i0]
                           : i < i1             ? cell_off[i] - (row_off[i] - d0)
                                                : 0;
    // CSR_U entries per lane at a time: their records, then their output
    // rows' offsets, are all requested before any is used.  The records are
    // read once, with nontemporal loads (through the caches they evicted the
    // output rows' offsets: the CSR phase 0.329 -> 0.318 ms at 2049^2 ->
    // 1537^2)
    constexpr int CSR_U = 4;
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    for (int64_t q00 = 0; q00 < cnt; q00 += 64 * CSR_U) {  // uniform trip count: shuffles below
        KV r[CSR_U];
        int32_t rw[CSR_U];
#pragma unroll
        for (int u = 0; u < CSR_U; u++) {
            const int64_t q = q00 + 64 * u + lane;
            // row of entry q = the last lane whose first entry is at or before q
            int32_t row = 0;
#pragma unroll
            for (int b = 32; b > 0; b >>= 1) {
                const int64_t dd = __shfl(my_d, (row + b) & 63, 64);
                if (row + b < 64 && dd <= q) row += b;
            }
            rw[u] = row;
            const int64_t src = __shfl(my_src, row, 64) + q;
            if (q < cnt) r[u] = __builtin_bit_cast(KV, __builtin_nontemporal_load((const u32x4*)&rec[src]));
        }
        int64_t tb[CSR_U];
#pragma unroll
        for (int u = 0; u < CSR_U; u++)
            if (q00 + 64 * u + lane < cnt) tb[u] = toff[(int32_t)(uint32_t)r[u].key];
#pragma unroll
        for (int u = 0; u < CSR_U; u++) {
            const int64_t q = q00 + 64 * u + lane;
            if (q < cnt) {
                const int32_t o = (int32_t)(uint32_t)r[u].key;
                col[d0 + q] = o;
                out_val[d0 + q] = r[u].val;
                const int64_t pos = tb[u] + (int64_t)(r[u].key >> 32);
                if (pos < nnz_cap) {
                    t_col[pos] = (int32_t)(i0 + rw[u] - P.cell_gbase);
                    t_val[pos] = r[u].val;
                }
            }
        }
    }
}


// Orders every row of the transposed CSR by column (the rows were filled
// in atomic order).  One wavefront per 64 consecutive rows, whose entries
// are one contiguous range: loaded into LDS with coalesced loads, sorted
// per lane in LDS, stored back coalesced.  Oversized groups sort in global
// memory.  Also writes each pair's int32 row pointers (pair k's n_out + 1
// pointers start at rowptr_all[out_gbase + k]).
constexpr int SORT_WAVES = 2;
constexpr int SORT_CAP = 1024;  // entries per wavefront (12 KiB)

// One transposed row (columns, values) sorted in registers through a
// network of N, written back in place.
template <int N>
__device__ inline void sort_row_in_registers(int32_t* C, double* V, int32_t n)
{
    int32_t k[N];
    double v[N];
#pragma unroll
    for (int j = 0; j < N; j++) {
        k[j] = INT32_MAX;  // padding sorts last (columns are below it)
        v[j] = 0;
        if (j < n) {
            k[j] = C[j];
            v[j] = V[j];
        }
    }
    sort_network<N, int32_t>(k, v);
#pragma unroll
    for (int j = 0; j < N; j++)
        if (j < n) {
            C[j] = k[j];
            V[j] = v[j];
        }
}

__global__ __launch_bounds__(SORT_WAVES * 64) void k_rows_sort(
    const PairCells* __restrict__ pcs, const int32_t* __restrict__ pair_of_group,
    const int64_t* __restrict__ off, int64_t nrows_all, int32_t* __restrict__ col,
    double* __restrict__ val, int32_t* __restrict__ rowptr_all, int64_t nnz_cap)
{
    __shared__ int32_t sc[SORT_WAVES][SORT_CAP];
    __shared__ double sv[SORT_WAVES][SORT_CAP];
    const int wid = block_wave(), lane = threadIdx.x & 63;
    const int64_t r0 = ((int64_t)blockIdx.x * SORT_WAVES + wid) * 64;
    if (r0 >= nrows_all) return;  // uniform per wavefront
    const int k = __builtin_amdgcn_readfirstlane(pair_of_group[r0 >> 6]);
    const PairCells& P = pcs[k];
    const int64_t rend = P.out_gbase + P.n_out;
    const int64_t obase = off[P.out_gbase];
    int32_t* rowptr = rowptr_all + P.out_gbase + k;
    const int64_t r = r0 + lane;
    if (r < rend) rowptr[r - P.out_gbase] = (int32_t)(off[r] - obase);
    if (r == rend - 1) rowptr[P.n_out] = (int32_t)(off[rend] - obase);
    const int64_t r1 = r0 + 64 < nrows_all ? r0 + 64 : nrows_all;
    if (off[r1] > nnz_cap) return;  // past a capacity-mode build's capacity
    const int64_t base = off[r0], total = off[r1] - base;
    int64_t s = 0, n = 0;
    if (r < r1) {
        s = off[r] - base;
        n = off[r + 1] - off[r];
    }
    if (total <= SORT_CAP) {
        int32_t* C = sc[wid];
        double* V = sv[wid];
#pragma unroll 4
        for (int64_t q = lane; q < total; q += 64) {
            C[q] = col[base + q];
            V[q] = val[base + q];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (r < r1) {
            if (n <= 8)
                sort_row_in_registers<8>(C + s, V + s, (int32_t)n);
            else if (n <= 16)
                sort_row_in_registers<16>(C + s, V + s, (int32_t)n);
            else
                sort_pairs(C + s, V + s, n);
        }
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

// Per pair: its first entry (row_off at its first cell) and its first
// triple (chunk_off at its first chunk); slot npairs holds the totals.
__global__ void k_pair_bases(const PairCells* __restrict__ pcs,
                             const Build* __restrict__ builds, int64_t npairs,
                             const int64_t* __restrict__ row_off, int64_t ncells_all,
                             const int64_t* __restrict__ chunk_off,
                             int64_t nchunks_all, const int32_t* __restrict__ counters,
                             int64_t* __restrict__ out)
{
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k > npairs) return;
    out[2 * k] = row_off[k < npairs ? pcs[k].cell_gbase : ncells_all];
    out[2 * k + 1] = chunk_off[k < npairs ? builds[k].chunk_gbase : nchunks_all];
    if (k == npairs) {  // the eight counters after the bases: one read for all
        int32_t* c = (int32_t*)(out + 2 * (npairs + 1));
        for (int q = 0; q < 8; q++) c[q] = counters[q];
    }
}

// The triple total, the histogram total and the eight counters in one
// buffer, read back with one copy.
__global__ void k_pack_totals(const int64_t* __restrict__ a, const int64_t* __restrict__ b,
                              const int32_t* __restrict__ counters, int64_t* __restrict__ out)
{
    out[0] = *a;
    out[1] = *b;
    int32_t* c = (int32_t*)(out + 2);
    for (int q = 0; q < 8; q++) c[q] = counters[q];
}

__global__ void k_fill_i32(int32_t* p, int32_t v, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ---- host orchestration --------------------------------------------------

namespace {

int fill_zero(int32_t* p, int64_t n, hipStream_t s)
{
    if (n == 0) return RG_OK;
    k_fill_i32<<<grid_for(n, 256) > 4096 ? 4096 : grid_for(n, 256), 256, 0,
                 s>>>(p, 0, n);
    RG_HIP(hipGetLastError());
    return RG_OK;
}

int launch_segments(const int64_t* prefix, int n, int64_t nunits, int32_t* out,
                    hipStream_t s)
{
    if (nunits == 0) return RG_OK;
    k_segments<<<grid_for(nunits, 256), 256, 0, s>>>(prefix, n, nunits, out);
    RG_HIP(hipGetLastError());
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

// Pinned host staging of the build's descriptor tables: one grow-only
// buffer per host thread.  The upload is asynchronous; every build waits on
// its stream (its first scalar read) before it returns, and an early error
// return waits here, so the next build of the thread may reuse the buffer.
struct PinnedUpload {
    hipStream_t s = nullptr;
    bool pending = false;
    ~PinnedUpload()
    {
        if (pending) (void)hipStreamSynchronize(s);
    }
    static int get(void** p, size_t bytes)
    {
        static thread_local void* buf = nullptr;
        static thread_local size_t cap = 0;
        if (bytes > cap) {
            if (buf) (void)hipHostFree(buf);
            buf = nullptr;
            cap = 0;
            const size_t sz = std::max<size_t>(bytes, 1 << 16);
            RG_HIP(hipHostMalloc(&buf, sz, hipHostMallocDefault));
            cap = sz;
        }
        *p = buf;
        return RG_OK;
    }
};

int64_t round64(int64_t n) { return (n + 63) / 64 * 64; }

template <class T>
size_t append_table(std::vector<unsigned char>& buf, const T* src, size_t n)
{
    size_t off = (buf.size() + 255) / 256 * 256;
    buf.resize(off + n * sizeof(T));
    if (n) std::memcpy(buf.data() + off, src, n * sizeof(T));
    return off;
}

}  // namespace

// The build of npairs independent grid pairs: every phase is one launch over
// the concatenation of all pairs' chunks, rows and (padded) cells.  A single
// pair is a batch of one.
// Capacity mode of build_batched, for capture in a hipGraph: every
// data-dependent size (triples, overflow triples, weights) is a fixed
// capacity, nothing is read back to the host, kernels that could pass a
// capacity skip their work and raise a bit in the flag word instead, and
// the build ends with the FP64 batched apply of one job per pair.
struct GraphJob {
    const double* x;  // (batch, n_in) values of the pair's input cells
    double* y;        // (batch, n_out)
    int64_t batch;    // 0: no apply for this pair
};

struct CaptureCaps {
    // false: an ordinary build that records its totals (T_seen ...) and
    // makes handles; true: the capacity mode described above
    bool capacity = false;
    int64_t T_seen = 0, ov_seen = 0, nnz_seen = 0;
    int64_t T_cap = 0, ov_cap = 0, nnz_cap = 0;
    void* meta_pinned = nullptr;  // caller-owned pinned snapshot of the tables
    size_t meta_pinned_cap = 0;
    const GraphJob* jobs = nullptr;  // one per pair, or null
    // filled by build_batched (device memory of the caller's arena)
    int32_t* counters = nullptr;  // word 8: capacity bits (CAP_*), words 0-7 as in a build
    int64_t* bases = nullptr;     // per pair: first weight, first triple (slot npairs: totals)
    int32_t *rp_in = nullptr, *c_in = nullptr, *rp_out = nullptr, *c_out = nullptr;
    double *v_in = nullptr, *v_out = nullptr;
    size_t meta_bytes = 0;        // bytes of the table snapshot
    std::vector<PairCells> pcs;
    std::vector<int64_t> nchunks;
};


__global__ void k_caps_check_triples(const int64_t* __restrict__ t_chunks,
                                     const int64_t* __restrict__ t_hist,
                                     int32_t* __restrict__ counters, int64_t T_cap, int64_t ov_cap)
{
    int32_t f = 0;
    if (*t_chunks > T_cap) f |= CAP_TRIPLES;
    if (*t_chunks != *t_hist) f |= CAP_MISMATCH;
    if (counters[2] != 0) f |= CAP_FAILED;
    int64_t ov;
    ov = (int64_t)((const unsigned long long*)(counters + 6))[0];
    if (ov > ov_cap) f |= CAP_OVERFLOW;
    if (f) atomicOr(&counters[8], f);
}

__global__ void k_caps_check_weights(const int64_t* __restrict__ nnz_total,
                                     int32_t* __restrict__ counters, int64_t nnz_cap)
{
    int32_t f = 0;
    if (*nnz_total > nnz_cap) f |= CAP_WEIGHTS;
    if (counters[2] != 0) f |= CAP_MISMATCH;
    if (f) atomicOr(&counters[8], f);
}

// The job table of the graph's apply: each pair's CSR by output sits at its
// first weight (known on the device only).
__global__ void k_graph_jobs(JobDesc* __restrict__ jobs, const int32_t* __restrict__ job_pair,
                             int njobs, const PairCells* __restrict__ pcs,
                             const int64_t* __restrict__ bases, const int32_t* __restrict__ rp_out,
                             const int32_t* __restrict__ c_out, const double* __restrict__ v_out)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= njobs) return;
    const int k = job_pair[j];
    jobs[j].rowptr = rp_out + pcs[k].out_gbase + k;
    jobs[j].col = c_out + bases[2 * k];
    jobs[j].w = v_out + bases[2 * k];
}

int build_batched(const rg_pair_desc* pairs, int64_t npairs,
                  const rg_tuning* tuning, rg_weights** out, hipStream_t s,
                  CaptureCaps* caps);

// The ordinary build (also called from the serving runtime).
int build_batched(const rg_pair_desc* pairs, int64_t npairs,
                  const rg_tuning* tuning, rg_weights** out, hipStream_t s)
{
    return build_batched(pairs, npairs, tuning, out, s, nullptr);
}

int build_batched(const rg_pair_desc* pairs, int64_t npairs,
                  const rg_tuning* tuning, rg_weights** out, hipStream_t s,
                  CaptureCaps* caps)
{
    const bool capmode = caps && caps->capacity;
    if (npairs <= 0) return fail(RG_ERR_ARGUMENT, "npairs must be positive");
    if (!pairs || (!out && !capmode)) return fail(RG_ERR_ARGUMENT, "pairs or out is NULL");
    if (out)
        for (int64_t k = 0; k < npairs; k++) out[k] = nullptr;
    if (injected_failure("build"))
        return fail(RG_ERR_INJECTED, "injected build failure");
    if (npairs >= (int64_t)1 << 24)
        return fail(RG_ERR_ARGUMENT, "too many pairs in one batch");
    for (int64_t k = 0; k < npairs; k++) {
        const rg_pair_desc& d = pairs[k];
        const std::string at = npairs > 1 ? " (pair " + std::to_string(k) + ")" : "";
        if (!d.xi || !d.yi || !d.xo || !d.yo)
            return fail(RG_ERR_ARGUMENT, "coordinate pointer is NULL" + at);
        if (d.nxi < 2 || d.nyi < 2 || d.nxo < 2 || d.nyo < 2)
            return fail(RG_ERR_ARGUMENT,
                        "each grid needs at least 2 vertices along both axes" + at);
        if (d.nxi * d.nyi >= INT32_MAX || d.nxo * d.nyo >= INT32_MAX)
            return fail(RG_ERR_ARGUMENT, "more than 2^31 - 1 vertices in one grid" + at);
    }
    int64_t K = 16, stage_per_edge = 8;
    int block = 128;
    bool timings = false;
    int32_t flags = 0;
    if (tuning) {
        if (tuning->edges_per_lane > 0) K = tuning->edges_per_lane;
        if (tuning->block > 0) block = tuning->block;
        if (tuning->stage_per_edge > 0) stage_per_edge = tuning->stage_per_edge;
        flags = tuning->flags;
        timings = (flags & RG_TUNE_TIMINGS) != 0 && !capmode;
    }
    if (block % 64 != 0 || block > WALK_MAX_BLOCK)
        return fail(RG_ERR_ARGUMENT, "block must be a multiple of 64, <= 512");

    PhaseTimer tm(timings, s);
    tm.mark();  // 0

    // ---- host tables: grids (2k input, 2k+1 output), cells, passes
    const int ng = (int)(2 * npairs);
    std::vector<GridDesc> grids(ng);
    std::vector<int64_t> tile_pf(ng + 1), e_pf(ng + 1), blk_pf(ng + 1), sup_pf(ng + 1),
        pyr_pf(ng + 1);
    int64_t v_tot = 0, e_tot = 0, blk_tot = 0, sup_tot = 0, tile_tot = 0, pyr_tot = 0;
    for (int g = 0; g < ng; g++) {
        const rg_pair_desc& d = pairs[g / 2];
        GridDesc& G = grids[g];
        G.x = g % 2 ? d.xo : d.xi;
        G.y = g % 2 ? d.yo : d.yi;
        G.nx = g % 2 ? d.nxo : d.nxi;
        G.ny = g % 2 ? d.nyo : d.nyi;
        const int64_t nv = G.nx * G.ny, nbe = 2 * (G.nx - 1) + 2 * (G.ny - 1);
        const int64_t nblk = (nbe + BND_BLOCK - 1) / BND_BLOCK;
        const int64_t nsup = (nblk + BND_SUPER - 1) / BND_SUPER;
        G.v_base = v_tot;
        G.e_base = e_pf[g] = e_tot;
        G.blk_base = blk_pf[g] = blk_tot;
        G.pyr_base = pyr_pf[g] = pyr_tot;
        pyr_tot += (nbe + 63) / 64;
        G.sup_base = sup_pf[g] = sup_tot;
        G.tile_base = tile_pf[g] = tile_tot;
        v_tot += nv;
        e_tot += nbe;
        blk_tot += nblk;
        sup_tot += nsup;
        tile_tot += (nv + VTILE - 1) / VTILE;
    }
    tile_pf[ng] = tile_tot;
    e_pf[ng] = e_tot;
    blk_pf[ng] = blk_tot;
    pyr_pf[ng] = pyr_tot;
    sup_pf[ng] = sup_tot;

    DevBuf<double2> xy_all;
    DevBuf<double> part, bbox_all, bpx, bpy, be;
    DevBuf<int32_t> bcell;
    DevBuf<Box> bblk, bsup, ball;
    DevBuf<double2> pyr;
    RG_TRY(xy_all.alloc(v_tot, s));
    RG_TRY(part.alloc(4 * tile_tot, s));
    RG_TRY(bbox_all.alloc(4 * ng, s));
    RG_TRY(bpx.alloc(e_tot, s));
    RG_TRY(bpy.alloc(e_tot, s));
    RG_TRY(be.alloc(4 * e_tot, s));
    RG_TRY(bcell.alloc(e_tot, s));
    RG_TRY(bblk.alloc(blk_tot, s));
    RG_TRY(ball.alloc(ng, s));
    RG_TRY(pyr.alloc(pyr_tot, s));
    RG_TRY(bsup.alloc(sup_tot, s));
    std::vector<StaticGrid> sg(ng);
    for (int g = 0; g < ng; g++) {
        const GridDesc& G = grids[g];
        StaticGrid& S = sg[g];
        S.x = G.x;
        S.y = G.y;
        S.xy = xy_all.p + G.v_base;
        S.nx = G.nx;
        S.ny = G.ny;
        S.bbox = bbox_all.p + 4 * g;
        S.bpx = bpx.p + G.e_base;
        S.bpy = bpy.p + G.e_base;
        S.nbp = e_pf[g + 1] - e_pf[g];
        S.be = be.p + 4 * G.e_base;
        S.bcell = bcell.p + G.e_base;
        S.nbe = S.nbp;
        S.bblk = bblk.p + G.blk_base;
        S.nblk = blk_pf[g + 1] - blk_pf[g];
        S.pyr = pyr.p + G.pyr_base;
        S.npyr = pyr_pf[g + 1] - pyr_pf[g];
        S.bsup = bsup.p + G.sup_base;
        S.ball = ball.p + g;
        S.nsup = sup_pf[g + 1] - sup_pf[g];
    }

    std::vector<PairCells> pcs(npairs);
    std::vector<Build> builds(npairs);
    std::vector<int64_t> wave_pf(npairs + 1), row_pf(npairs + 1),
        ing_pf(npairs + 1), outg_pf(npairs + 1);
    int64_t cell_tot = 0, out_tot = 0, thr_tot = 0, chunk_tot = 0, row_tot = 0;
    bool any_wi = false;
    for (int64_t k = 0; k < npairs; k++) {
        const rg_pair_desc& d = pairs[k];
        PairCells& P = pcs[k];
        P.xi = d.xi;
        P.yi = d.yi;
        P.xyi = sg[2 * k].xy;
        P.nyi = d.nyi;
        P.wi = d.wi;
        any_wi = any_wi || d.wi;
        P.n_in = (d.nxi - 1) * (d.nyi - 1);
        P.n_out = (d.nxo - 1) * (d.nyo - 1);
        P.cell_gbase = cell_tot;
        P.out_gbase = out_tot;
        Build& B = builds[k];
        B.K = K;
        int64_t chunk = 0, rows = 0, threads = 0;
        for (int p = 0; p < 4; p++) {
            Pass& Q = B.pass[p];
            const int sweep_input = p >= 2;
            const int axis = p & 1;
            const int64_t nxs = sweep_input ? d.nxi : d.nxo;
            const int64_t nys = sweep_input ? d.nyi : d.nyo;
            Q.xs = sweep_input ? d.xi : d.xo;
            Q.ys = sweep_input ? d.yi : d.yo;
            Q.xys = sweep_input ? sg[2 * k].xy : sg[2 * k + 1].xy;
            if (axis == 0) {  // aligned view = transpose
                Q.sr = 1; Q.sk = nys; Q.nrows = nys; Q.nk = nxs;
            } else {
                Q.sr = nys; Q.sk = 1; Q.nrows = nxs; Q.nk = nys;
            }
            Q.ncy_sweep = nys - 1;
            Q.axis = axis;
            Q.sweep_input = sweep_input;
            Q.st = sweep_input ? sg[2 * k + 1] : sg[2 * k];
            Q.cpr = std::max<int64_t>(1, (Q.nk - 1 + K - 1) / K);
            Q.chunk_base = chunk;
            Q.thread_base = threads;
            B.row_base[p] = rows;
            chunk += Q.nrows * Q.cpr;
            rows += Q.nrows;
            threads += threads_of_pass(Q.nrows, Q.cpr);
        }
        B.nchunks = chunk;
        B.nrows_total = rows;
        B.nthreads = threads;
        B.thread_gbase = thr_tot;
        B.chunk_gbase = chunk_tot;
        B.row_gbase = row_tot;
        B.cell_gbase = cell_tot;
        B.out_gbase = out_tot;
        wave_pf[k] = thr_tot / 64;
        row_pf[k] = row_tot;
        ing_pf[k] = cell_tot / 64;
        outg_pf[k] = out_tot / 64;
        thr_tot += threads;
        chunk_tot += chunk;
        row_tot += rows;
        cell_tot += round64(P.n_in);
        out_tot += round64(P.n_out);
    }
    wave_pf[npairs] = thr_tot / 64;
    row_pf[npairs] = row_tot;
    ing_pf[npairs] = cell_tot / 64;
    outg_pf[npairs] = out_tot / 64;
    if (thr_tot >= INT32_MAX || thr_tot % 64 != 0)
        return fail(RG_ERR_ARGUMENT, "too many sweep chunks in one build");

    // ---- upload the tables (one copy)
    std::vector<unsigned char> hb;
    const size_t o_builds = append_table(hb, builds.data(), builds.size());
    const size_t o_grids = append_table(hb, grids.data(), grids.size());
    const size_t o_pcs = append_table(hb, pcs.data(), pcs.size());
    const size_t o_tile = append_table(hb, tile_pf.data(), tile_pf.size());
    const size_t o_e = append_table(hb, e_pf.data(), e_pf.size());
    const size_t o_blk = append_table(hb, blk_pf.data(), blk_pf.size());
    const size_t o_pyr = append_table(hb, pyr_pf.data(), pyr_pf.size());
    const size_t o_sup = append_table(hb, sup_pf.data(), sup_pf.size());
    const size_t o_wave = append_table(hb, wave_pf.data(), wave_pf.size());
    const size_t o_row = append_table(hb, row_pf.data(), row_pf.size());
    const size_t o_ing = append_table(hb, ing_pf.data(), ing_pf.size());
    const size_t o_outg = append_table(hb, outg_pf.data(), outg_pf.size());
    // the graph's apply jobs (shapes only; the CSR pointers are set on the
    // device once the weights exist)
    std::vector<JobDesc> gjobs;
    std::vector<int64_t> gprefix;
    std::vector<int32_t> gpair;
    int64_t gblocks = 0;
    if (capmode && caps->jobs) {
        for (int64_t k = 0; k < npairs; k++) {
            const GraphJob& J = caps->jobs[k];
            if (J.batch <= 0 || pcs[k].n_out == 0) continue;
            JobDesc d = {};
            d.x = J.x;
            d.y = J.y;
            d.nrows = pcs[k].n_out;
            d.ncols = pcs[k].n_in;
            d.batch = J.batch;
            job_tiles(d, pairs[k].nxo - 1, pairs[k].nyo - 1, 8);
            d.block_base = gblocks;
            gprefix.push_back(gblocks);
            gblocks += d.nrt * ((d.batch + JOB_DB - 1) / JOB_DB);
            gjobs.push_back(d);
            gpair.push_back((int32_t)k);
        }
    }
    const size_t o_gjobs = append_table(hb, gjobs.data(), gjobs.size());
    const size_t o_gprefix = append_table(hb, gprefix.data(), gprefix.size());
    const size_t o_gpair = append_table(hb, gpair.data(), gpair.size());
    DevBuf<unsigned char> meta;
    RG_TRY(meta.alloc(hb.size(), s));
    PinnedUpload up;
    up.s = s;
    void* pinned = nullptr;
    if (capmode) {
        // a snapshot the graph keeps: its copy node re-reads it every replay
        if (!caps->meta_pinned) {  // first (uncaptured) run: make it
            RG_HIP(hipHostMalloc(&caps->meta_pinned, hb.size(), hipHostMallocDefault));
            caps->meta_pinned_cap = hb.size();
        }
        if (hb.size() > caps->meta_pinned_cap)
            return fail(RG_ERR_ARGUMENT, "internal: graph table snapshot too small");
        pinned = caps->meta_pinned;
        caps->meta_bytes = hb.size();
    } else {
        RG_TRY(PinnedUpload::get(&pinned, hb.size()));
        up.pending = true;
    }
    std::memcpy(pinned, hb.data(), hb.size());
    RG_HIP(hipMemcpyAsync(meta.p, pinned, hb.size(), hipMemcpyHostToDevice, s));
    const Build* d_builds = (const Build*)(meta.p + o_builds);
    const GridDesc* d_grids = (const GridDesc*)(meta.p + o_grids);
    const PairCells* d_pcs = (const PairCells*)(meta.p + o_pcs);
    const int64_t* d_tile = (const int64_t*)(meta.p + o_tile);
    const int64_t* d_e = (const int64_t*)(meta.p + o_e);
    const int64_t* d_blk = (const int64_t*)(meta.p + o_blk);
    const int64_t* d_pyr = (const int64_t*)(meta.p + o_pyr);
    const int64_t* d_sup = (const int64_t*)(meta.p + o_sup);

    // ---- prepare
    DevBuf<int32_t> pair_of_wave, pair_of_row, pair_of_ing, pair_of_outg;
    RG_TRY(pair_of_wave.alloc(thr_tot / 64, s));
    RG_TRY(pair_of_row.alloc(row_tot, s));
    RG_TRY(pair_of_ing.alloc(cell_tot / 64, s));
    RG_TRY(pair_of_outg.alloc(out_tot / 64, s));
    const int np32 = (int)npairs;
    RG_TRY(launch_segments((const int64_t*)(meta.p + o_wave), np32, thr_tot / 64,
                           pair_of_wave.p, s));
    RG_TRY(launch_segments((const int64_t*)(meta.p + o_row), np32, row_tot,
                           pair_of_row.p, s));
    RG_TRY(launch_segments((const int64_t*)(meta.p + o_ing), np32, cell_tot / 64,
                           pair_of_ing.p, s));
    RG_TRY(launch_segments((const int64_t*)(meta.p + o_outg), np32, out_tot / 64,
                           pair_of_outg.p, s));
    k_grid_tiles<<<(unsigned)tile_tot, BBOX_BLOCK, 0, s>>>(d_grids, d_tile, ng,
                                                            xy_all.p, part.p);
    RG_HIP(hipGetLastError());
    k_grid_bbox<<<(unsigned)ng, BBOX_BLOCK, 0, s>>>(d_tile, part.p, bbox_all.p);
    RG_HIP(hipGetLastError());
    k_grid_boundary<<<grid_for(e_tot, 256), 256, 0, s>>>(d_grids, d_e, ng, e_tot,
                                                         bpx.p, bpy.p, be.p, bcell.p);
    RG_HIP(hipGetLastError());
    k_grid_edge_boxes<<<grid_for(blk_tot, 256), 256, 0, s>>>(d_grids, d_blk, ng,
                                                             blk_tot, be.p, bblk.p);
    RG_HIP(hipGetLastError());
    k_grid_poly_yranges<<<grid_for(pyr_tot * 64, 256), 256, 0, s>>>(d_grids, d_pyr, ng,
                                                                    pyr_tot, bpy.p, pyr.p);
    RG_HIP(hipGetLastError());
    k_grid_super_boxes<<<grid_for(sup_tot, 256), 256, 0, s>>>(d_grids, d_sup, ng,
                                                              sup_tot, bblk.p, bsup.p);
    RG_HIP(hipGetLastError());
    k_grid_boundary_box<<<grid_for(ng, 64), 64, 0, s>>>(d_sup, ng, bsup.p, ball.p);
    RG_HIP(hipGetLastError());
    DevBuf<double> vol, wi_all;
    RG_TRY(vol.alloc(cell_tot, s));
    if (any_wi) RG_TRY(wi_all.alloc(cell_tot, s));
    k_pair_cells<<<grid_for(cell_tot, 256), 256, 0, s>>>(d_pcs, pair_of_ing.p,
                                                         cell_tot, vol.p, wi_all.p);
    RG_HIP(hipGetLastError());
    tm.mark();  // 1

    // ---- start states
    const int64_t n_groups = cell_tot / 64;  // every pair's cells padded to 64
    if (4 * n_groups >= INT32_MAX) return fail(RG_ERR_ARGUMENT, "too many input cells in one build");
    DevBuf<int32_t> start, end, count, hist, counters;
    RG_TRY(start.alloc(chunk_tot, s));
    RG_TRY(end.alloc(chunk_tot, s));
    RG_TRY(count.alloc(chunk_tot, s));
    RG_TRY(hist.alloc(4 * n_groups, s));
    // anomalies, fixups, failed/overflow, long runs, restaged chunks, heavy
    // groups; words 6-7: the triples beyond the staging capacity (64-bit);
    // word 8: capacity bits of a capacity-mode build; word 9: the heavy
    // groups' work-queue cursor
    RG_TRY(counters.alloc(12, s));
    RG_TRY(fill_zero(hist.p, 4 * n_groups, s));
    RG_TRY(fill_zero(counters.p, 12, s));
    {
        // rows per wavefront: one until the rows fill about 4096 wavefronts
        int rpw = 1;
        while (rpw < 64 && row_tot / (2 * rpw) >= 4096) rpw *= 2;
        if (flags & RG_TUNE_ROW_WAVES) rpw = 64;
        const int64_t waves = (row_tot + rpw - 1) / rpw;
        k_row_start<<<grid_for(waves * 64, 256), 256, 0, s>>>(d_builds, pair_of_row.p, row_tot,
                                                              start.p, counters.p + 0, rpw);
    }
    RG_HIP(hipGetLastError());
    k_chunk_start<<<grid_for(thr_tot, block), block, 0, s>>>(
        d_builds, pair_of_wave.p, thr_tot, start.p, (flags & RG_TUNE_GUESS_OUTSIDE) != 0);
    RG_HIP(hipGetLastError());
    tm.mark();  // 2

    // ---- count (staging the triples)
    const int64_t cap = K * stage_per_edge;
    DevBuf<Triple> stage;
    RG_TRY(stage.alloc((size_t)thr_tot * cap, s));
    k_walk_count<<<grid_for(thr_tot, block), block, 0, s>>>(
        d_builds, pair_of_wave.p, thr_tot, n_groups, start.p, end.p, count.p, hist.p,
        stage.p, cap);
    RG_HIP(hipGetLastError());
    tm.mark();  // 3

    // ---- fix-up
    k_fixup<<<grid_for(row_tot * 64, 256), 256, 0, s>>>(
        d_builds, pair_of_row.p, row_tot, n_groups, start.p, end.p, count.p, hist.p,
        counters.p + 1, counters.p + 2, stage.p, cap);
    RG_HIP(hipGetLastError());
    DevBuf<int64_t> chunk_off, bucket_off;
    RG_TRY(chunk_off.alloc(chunk_tot + 1, s));
    RG_TRY(bucket_off.alloc(4 * n_groups + 1, s));
    RG_TRY(scan_i32_to_i64(count.p, chunk_off.p, chunk_tot, s));
    RG_TRY(scan_i32_to_i64(hist.p, bucket_off.p, 4 * n_groups, s));
    // triples beyond the staging capacity, per chunk and in all
    DevBuf<int32_t> ov;
    RG_TRY(ov.alloc(chunk_tot, s));
    k_overflow_counts<<<grid_for(chunk_tot, 256), 256, 0, s>>>(
        count.p, chunk_tot, cap, ov.p, (unsigned long long*)(counters.p + 6));
    RG_HIP(hipGetLastError());
    int64_t tot[2];
    int32_t cnt[8] = {};
    int64_t ov_tot;
    int64_t T;
    if (capmode) {
        k_caps_check_triples<<<1, 1, 0, s>>>(chunk_off.p + chunk_tot, bucket_off.p + 4 * n_groups,
                                             counters.p, caps->T_cap, caps->ov_cap);
        RG_HIP(hipGetLastError());
        T = caps->T_cap;
        ov_tot = caps->ov_cap;
    } else {
        DevBuf<int64_t> packed;
        RG_TRY(packed.alloc(6, s));
        k_pack_totals<<<1, 1, 0, s>>>(chunk_off.p + chunk_tot, bucket_off.p + 4 * n_groups,
                                      counters.p, packed.p);
        RG_HIP(hipGetLastError());
        int64_t hp[6];
        RG_TRY(read_scalars(packed.p, hp, sizeof(hp), s));
        tot[0] = hp[0];
        tot[1] = hp[1];
        std::memcpy(cnt, hp + 2, sizeof(cnt));
        std::memcpy(&ov_tot, cnt + 6, sizeof(ov_tot));
        up.pending = false;  // the stream has passed the upload
        if (cnt[2] != 0)
            return fail(RG_ERR_GEOMETRY,
                        "the sweep did not terminate in " + std::to_string(cnt[2]) +
                            " rows (exactly degenerate crossings)");
        T = tot[0];
        if (T != tot[1])
            return fail(RG_ERR_HIP, "internal: triple count " + std::to_string(T) +
                                        " != histogram total " +
                                        std::to_string(tot[1]));
        if (caps) {
            caps->T_seen = T;
            caps->ov_seen = ov_tot;
        }
    }
    tm.mark();  // 4
    if (T >= ((int64_t)1 << 32))
        return fail(RG_ERR_ARGUMENT, "more than 2^32 triples in one build");

    // ---- emit
    DevBuf<KV> rec;
    DevBuf<uint8_t> rcell;
    RG_TRY(rec.alloc(T, s));
    RG_TRY(rcell.alloc(T, s));
    uint32_t* cursor = (uint32_t*)hist.p;  // reused: one word per bucket
    k_cursor_init<<<grid_for(4 * n_groups, 256), 256, 0, s>>>(bucket_off.p, cursor,
                                                              4 * n_groups);
    tm.mark();  // 5
    DevBuf<Triple> ovbuf;
    DevBuf<int64_t> ov_off;
    RG_TRY(ovbuf.alloc(ov_tot > 0 ? ov_tot : 1, s));
    RG_TRY(ov_off.alloc(chunk_tot + 1, s));
    if (ov_tot > 0 || capmode) {
        // the chunks' ranges of the overflow buffer
        RG_TRY(scan_i32_to_i64(ov.p, ov_off.p, chunk_tot, s));
        k_restage<<<grid_for(thr_tot, block), block, 0, s>>>(
            d_builds, pair_of_wave.p, thr_tot, start.p, count.p, ov_off.p, ovbuf.p, cap,
            ov_tot, counters.p + 4);
        RG_HIP(hipGetLastError());
    }
    k_scatter<<<grid_for(thr_tot, SCAT_BLOCK), SCAT_BLOCK, (SCAT_BLOCK / 64) * sizeof(ScatTable),
                s>>>(
        d_builds, pair_of_wave.p, thr_tot, n_groups, count.p, stage.p, cap, ovbuf.p,
        ov_off.p, ov_tot, chunk_off.p, cursor, rec.p, rcell.p, (uint32_t)T, counters.p + 2);
    RG_HIP(hipGetLastError());
    tm.mark();  // 6
    k_cursor_check<<<grid_for(4 * n_groups, 256), 256, 0, s>>>(
        bucket_off.p, cursor, 4 * n_groups, counters.p + 2);
    RG_HIP(hipGetLastError());
    stage.reset();

    // ---- reduce -> CSR by input
    DevBuf<int32_t> nnz_row, long_list, group_mode;
    DevBuf<int64_t> row_off, cell_off, group_off;
    DevBuf<KV> red;
    RG_TRY(nnz_row.alloc(cell_tot, s));
    RG_TRY(long_list.alloc(cell_tot, s));
    RG_TRY(group_mode.alloc(n_groups, s));
    RG_TRY(row_off.alloc(cell_tot + 1, s));
    RG_TRY(cell_off.alloc(cell_tot + 1, s));
    RG_TRY(group_off.alloc(n_groups + 1, s));
    RG_TRY(red.alloc(T, s));
    // group-major offsets of the reduction's output (group_mode as scratch)
    k_group_counts<<<grid_for(n_groups, 256), 256, 0, s>>>(bucket_off.p, n_groups,
                                                           group_mode.p);
    RG_HIP(hipGetLastError());
    RG_TRY(scan_i32_to_i64(group_mode.p, group_off.p, n_groups, s));
    tm.mark();  // 7
    DevBuf<int32_t> heavy_list, hist_out;
    RG_TRY(heavy_list.alloc(n_groups, s));
    // entries per output row (the reductions also keep each entry's rank)
    RG_TRY(hist_out.alloc(out_tot, s));
    RG_TRY(fill_zero(hist_out.p, out_tot, s));
    k_reduce_groups<<<(unsigned)((n_groups + RED_WAVES - 1) / RED_WAVES),
                      RED_WAVES * 64, 0, s>>>(bucket_off.p, n_groups, group_off.p, rec.p,
                                              rcell.p, red.p, cell_off.p, vol.p, wi_all.p,
                                              nnz_row.p, group_mode.p, long_list.p,
                                              counters.p + 3, heavy_list.p, counters.p + 5,
                                              d_pcs, pair_of_ing.p, hist_out.p, T,
                                              counters.p + 8);
    // (resident waves: four blocks of two waves per CU fit the LDS)
    k_reduce_heavy<<<1024, HEAVY_WAVES * 64, 0, s>>>(
        bucket_off.p, n_groups, group_off.p, rec.p, rcell.p, red.p, cell_off.p, vol.p,
        wi_all.p, nnz_row.p, group_mode.p, long_list.p, counters.p + 3, heavy_list.p,
        counters.p + 5, counters.p + 9, d_pcs, pair_of_ing.p, hist_out.p);
    k_bucket_long<<<256, 64, 0, s>>>(cell_off.p, red.p, vol.p, wi_all.p, long_list.p,
                                   counters.p + 3, nnz_row.p, d_pcs, pair_of_ing.p,
                                   hist_out.p);
    RG_HIP(hipGetLastError());
    tm.mark();  // 8
    RG_TRY(scan_i32_to_i64(nnz_row.p, row_off.p, cell_tot, s));
    // per-pair first entry and first triple (slot npairs: totals)
    DevBuf<int64_t> bases;
    RG_TRY(bases.alloc(2 * (npairs + 1) + 4, s));
    k_pair_bases<<<grid_for(npairs + 1, 256), 256, 0, s>>>(
        d_pcs, d_builds, npairs, row_off.p, cell_tot, chunk_off.p, chunk_tot, counters.p,
        bases.p);
    RG_HIP(hipGetLastError());
    std::vector<int64_t> hbases(2 * (npairs + 1) + 4);
    int32_t cnt2[8] = {};
    int64_t nnz;
    DevBuf<int64_t> off_out;
    if (capmode) {
        k_caps_check_weights<<<1, 1, 0, s>>>(row_off.p + cell_tot, counters.p, caps->nnz_cap);
        RG_HIP(hipGetLastError());
        nnz = caps->nnz_cap;
    } else {
        // the output rows' scan does not depend on the weight count: it runs
        // while the host waits for the count
        RG_TRY(read_scalars_begin(bases.p, sizeof(int64_t) * hbases.size(), s));
        RG_TRY(off_out.alloc(out_tot + 1, s));
        RG_TRY(scan_i32_to_i64(hist_out.p, off_out.p, out_tot, s));
        RG_TRY(read_scalars_end(hbases.data(), sizeof(int64_t) * hbases.size()));
        std::memcpy(cnt2, hbases.data() + 2 * (npairs + 1), sizeof(cnt2));
        if (cnt2[2] != 0)
            return fail(RG_ERR_HIP, "internal: the emission did not match the "
                                    "counted buckets (" + std::to_string(cnt2[2]) +
                                        " mismatches)");
        nnz = hbases[2 * npairs];
        if (caps) caps->nnz_seen = nnz;
        for (int64_t k = 0; k < npairs; k++)
            if (hbases[2 * k + 2] - hbases[2 * k] >= INT32_MAX)
                return fail(RG_ERR_ARGUMENT, "more than 2^31 - 1 weights in one element");
    }

    // ---- CSR by input, transpose -> CSR by output
    DevBuf<int32_t> rp_in, c_in, rp_out, c_out;
    DevBuf<double> v_in, v_out;
    RG_TRY(rp_in.alloc(cell_tot + npairs, s));
    RG_TRY(c_in.alloc(nnz, s));
    RG_TRY(v_in.alloc(nnz, s));
    RG_TRY(rp_out.alloc(out_tot + npairs, s));
    RG_TRY(c_out.alloc(nnz, s));
    RG_TRY(v_out.alloc(nnz, s));
    rec.reset();
    if (capmode) {
        RG_TRY(off_out.alloc(out_tot + 1, s));
        RG_TRY(scan_i32_to_i64(hist_out.p, off_out.p, out_tot, s));
    }
    k_csr_groups<<<grid_for(n_groups * 64, 256), 256, 0, s>>>(
        d_pcs, pair_of_ing.p, n_groups, cell_off.p, row_off.p, group_mode.p,
        red.p, rp_in.p, c_in.p, v_in.p, off_out.p, c_out.p, v_out.p, nnz);
    RG_HIP(hipGetLastError());
    const int64_t sort_groups = out_tot / 64;
    k_rows_sort<<<(unsigned)((sort_groups + SORT_WAVES - 1) / SORT_WAVES),
                  SORT_WAVES * 64, 0, s>>>(d_pcs, pair_of_outg.p, off_out.p, out_tot,
                                           c_out.p, v_out.p, rp_out.p, nnz);
    RG_HIP(hipGetLastError());
    tm.mark();  // 9

    if (capmode) {
        // the graph's apply, then hand the device arrays to the graph (they
        // live in its arena)
        if (!gjobs.empty()) {
            JobDesc* d_gjobs = (JobDesc*)(meta.p + o_gjobs);
            k_graph_jobs<<<grid_for((int64_t)gjobs.size(), 64), 64, 0, s>>>(
                d_gjobs, (const int32_t*)(meta.p + o_gpair), (int)gjobs.size(), d_pcs, bases.p,
                rp_out.p, c_out.p, v_out.p);
            RG_HIP(hipGetLastError());
            RG_TRY(apply_jobs_device(d_gjobs, (const int64_t*)(meta.p + o_gprefix),
                                     (int)gjobs.size(), gblocks, counters.p + 8, s));
        }
        caps->counters = counters.release();
        caps->bases = bases.release();
        caps->rp_in = rp_in.release();
        caps->c_in = c_in.release();
        caps->v_in = v_in.release();
        caps->rp_out = rp_out.release();
        caps->c_out = c_out.release();
        caps->v_out = v_out.release();
        caps->pcs = pcs;
        caps->nchunks.resize(npairs);
        for (int64_t k = 0; k < npairs; k++) caps->nchunks[k] = builds[k].nchunks;
        return RG_OK;
    }

    // ---- handles
    std::vector<rg_weights*> made;
    auto arena = std::make_shared<rg_arena>();
    int32_t* a_rp_in = rp_in.release();
    int32_t* a_c_in = c_in.release();
    double* a_v_in = v_in.release();
    int32_t* a_rp_out = rp_out.release();
    int32_t* a_c_out = c_out.release();
    double* a_v_out = v_out.release();
    arena->bufs = {a_rp_in, a_c_in, a_v_in, a_rp_out, a_c_out, a_v_out};
    int device = 0;
    (void)hipGetDevice(&device);
    rg_build_stats batch = {};
    batch.n_fixups = cnt[1];
    batch.n_search_anomalies = cnt[0];
    batch.n_restaged = cnt2[4];
    batch.n_long_cells = cnt2[3];
    if (timings) {
        batch.ms_prepare = tm.ms(0, 1);
        batch.ms_start = tm.ms(1, 2);
        batch.ms_count = tm.ms(2, 3);
        batch.ms_fixup = tm.ms(3, 4);
        batch.ms_emit = tm.ms(5, 6);
        batch.ms_reduce = tm.ms(7, 8);
        batch.ms_transpose = tm.ms(8, 9);
        batch.ms_total = tm.ms(0, 9);
    }
    for (int64_t k = 0; k < npairs; k++) {
        const rg_pair_desc& d = pairs[k];
        const PairCells& P = pcs[k];
        const int64_t eb = hbases[2 * k];
        rg_weights* W = new rg_weights();
        W->device = device;
        W->home = s;
        W->n_in = P.n_in;
        W->n_out = P.n_out;
        W->nnz = hbases[2 * k + 2] - eb;
        W->shape_in[0] = d.nxi - 1;
        W->shape_in[1] = d.nyi - 1;
        W->shape_out[0] = d.nxo - 1;
        W->shape_out[1] = d.nyo - 1;
        W->rowptr_in = a_rp_in + P.cell_gbase + k;
        W->col_in = a_c_in + eb;
        W->val_in = a_v_in + eb;
        W->rowptr_out = a_rp_out + P.out_gbase + k;
        W->col_out = a_c_out + eb;
        W->val_out = a_v_out + eb;
        W->arena = arena;
        rg_build_stats& S = W->stats;
        S = batch;
        S.n_triples = hbases[2 * k + 3] - hbases[2 * k + 1];
        S.n_chunks = builds[k].nchunks;
        // every inside sub-segment emits two triples except on the first and
        // last row of a sweep, which emit one
        S.n_subsegments = S.n_triples / 2;
        out[k] = W;
    }
    return RG_OK;
}

int build_conservative_2d(const double* xi, const double* yi, int64_t nxi,
                          int64_t nyi, const double* xo, const double* yo,
                          int64_t nxo, int64_t nyo, const double* wi,
                          const rg_tuning* tuning, rg_weights** out,
                          hipStream_t s)
{
    rg_pair_desc d;
    d.xi = xi;
    d.yi = yi;
    d.nxi = nxi;
    d.nyi = nyi;
    d.xo = xo;
    d.yo = yo;
    d.nxo = nxo;
    d.nyo = nyo;
    d.wi = wi;
    if (!out) return fail(RG_ERR_ARGUMENT, "out is NULL");
    return build_batched(&d, 1, tuning, out, s);
}

// ---- hipGraph of a batched build + apply ------------------------------

void weights_release(rg_weights* W);

struct GraphImpl {
    int device = 0;
    int64_t npairs = 0;
    std::vector<rg_pair_desc> pairs;
    CaptureCaps caps;
    void* arena = nullptr;
    size_t arena_bytes = 0;
    hipStream_t cs = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<GraphJob> jobs;
    bool launched = false;  // the buffers hold a replay's results

    ~GraphImpl()
    {
        (void)hipSetDevice(device);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (cs) {
            (void)hipStreamSynchronize(cs);
            (void)hipStreamDestroy(cs);
        }
        (void)hipDeviceSynchronize();  // replays on other streams
        if (arena) (void)hipFree(arena);
        if (caps.meta_pinned) (void)hipHostFree(caps.meta_pinned);
    }
};

namespace {

int64_t with_slack(int64_t n, double slack, int64_t floor_add)
{
    return (int64_t)((double)n * (1.0 + slack)) + floor_add;
}

// A copy of pair k's weights out of the graph's buffers.
int graph_weights_copy(GraphImpl* G, int64_t k, rg_weights** out, hipStream_t s)
{
    if (!G->launched) return fail(RG_ERR_ARGUMENT, "the graph has not been launched");
    int32_t flags = 0;
    RG_TRY(read_scalars(G->caps.counters + 8, &flags, sizeof(flags), s));
    if (flags)
        return fail(RG_ERR_CAPACITY, "the graph's last replay exceeded its capacity (flags " +
                                         std::to_string(flags) + ")");
    std::vector<int64_t> hb(2 * (G->npairs + 1));
    RG_TRY(read_scalars(G->caps.bases, hb.data(), sizeof(int64_t) * hb.size(), s));
    if (hb[2 * G->npairs] > G->caps.nnz_cap || hb[2 * k] > hb[2 * k + 2])
        return fail(RG_ERR_HIP, "internal: graph weight offsets out of range");
    const PairCells& P = G->caps.pcs[k];
    const rg_pair_desc& d = G->pairs[k];
    const int64_t eb = hb[2 * k], nnz = hb[2 * k + 2] - eb;
    DevBuf<int32_t> rp_in, c_in, rp_out, c_out;
    DevBuf<double> v_in, v_out;
    RG_TRY(rp_in.alloc(P.n_in + 1, s));
    RG_TRY(rp_out.alloc(P.n_out + 1, s));
    RG_TRY(c_in.alloc(nnz > 0 ? nnz : 1, s));
    RG_TRY(c_out.alloc(nnz > 0 ? nnz : 1, s));
    RG_TRY(v_in.alloc(nnz > 0 ? nnz : 1, s));
    RG_TRY(v_out.alloc(nnz > 0 ? nnz : 1, s));
    auto cp = [&](void* dst, const void* src, size_t bytes) -> int {
        if (bytes) RG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s));
        return RG_OK;
    };
    RG_TRY(cp(rp_in.p, G->caps.rp_in + P.cell_gbase + k, 4 * (P.n_in + 1)));
    RG_TRY(cp(rp_out.p, G->caps.rp_out + P.out_gbase + k, 4 * (P.n_out + 1)));
    RG_TRY(cp(c_in.p, G->caps.c_in + eb, 4 * nnz));
    RG_TRY(cp(c_out.p, G->caps.c_out + eb, 4 * nnz));
    RG_TRY(cp(v_in.p, G->caps.v_in + eb, 8 * nnz));
    RG_TRY(cp(v_out.p, G->caps.v_out + eb, 8 * nnz));
    RG_HIP(hipStreamSynchronize(s));
    rg_weights* W = new rg_weights();
    W->device = G->device;
    W->home = s;
    W->n_in = P.n_in;
    W->n_out = P.n_out;
    W->nnz = nnz;
    W->shape_in[0] = d.nxi - 1;
    W->shape_in[1] = d.nyi - 1;
    W->shape_out[0] = d.nxo - 1;
    W->shape_out[1] = d.nyo - 1;
    auto arena = std::make_shared<rg_arena>();
    W->rowptr_in = rp_in.release();
    W->col_in = c_in.release();
    W->val_in = v_in.release();
    W->rowptr_out = rp_out.release();
    W->col_out = c_out.release();
    W->val_out = v_out.release();
    arena->bufs = {W->rowptr_in, W->col_in, W->val_in, W->rowptr_out, W->col_out, W->val_out};
    W->arena = arena;
    W->stats.n_triples = hb[2 * k + 3] - hb[2 * k + 1];
    W->stats.n_subsegments = W->stats.n_triples / 2;
    W->stats.n_chunks = G->caps.nchunks[k];
    *out = W;
    return RG_OK;
}

}  // namespace

int graph_create(const rg_pair_desc* pairs, int64_t npairs, const rg_tuning* tuning,
                 const GraphJob* jobs, double slack, GraphImpl** out, hipStream_t s)
{
    *out = nullptr;
    if (!(slack >= -0.99 && slack <= 100)) return fail(RG_ERR_ARGUMENT, "slack out of range");
    if (npairs <= 0 || !pairs) return fail(RG_ERR_ARGUMENT, "no pairs");
    auto G = std::make_unique<GraphImpl>();
    RG_HIP(hipGetDevice(&G->device));
    G->npairs = npairs;
    G->pairs.assign(pairs, pairs + npairs);
    if (jobs) G->jobs.assign(jobs, jobs + npairs);
    rg_tuning t = {};
    if (tuning) t = *tuning;
    t.flags &= ~RG_TUNE_TIMINGS;
    // 1. an ordinary build measures the data-dependent sizes
    {
        std::vector<rg_weights*> hs(npairs, nullptr);
        CaptureCaps sizing;
        const int st = build_batched(pairs, npairs, &t, hs.data(), s, &sizing);
        for (rg_weights* w : hs) weights_release(w);
        if (st != RG_OK) return st;
        G->caps.T_cap = with_slack(sizing.T_seen, slack, 4096);
        G->caps.ov_cap = with_slack(sizing.ov_seen, slack, 4096);
        G->caps.nnz_cap = with_slack(sizing.nnz_seen, slack, 4096);
        if (G->caps.T_cap >= ((int64_t)1 << 32))
            return fail(RG_ERR_ARGUMENT, "more than 2^32 triples in one graph");
    }
    G->caps.capacity = true;
    G->caps.jobs = jobs ? G->jobs.data() : nullptr;
    // 2. a capacity-mode run, not captured, adds up the workspace it needs
    //    (and makes the table snapshot)
    {
        arena_measure_begin();
        const int st = build_batched(pairs, npairs, &t, nullptr, s, &G->caps);
        G->arena_bytes = arena_measure_end() + (1 << 20);
        if (st != RG_OK) return st;
        RG_HIP(hipStreamSynchronize(s));
        for (void* p : {(void*)G->caps.counters, (void*)G->caps.bases, (void*)G->caps.rp_in,
                        (void*)G->caps.c_in, (void*)G->caps.v_in, (void*)G->caps.rp_out,
                        (void*)G->caps.c_out, (void*)G->caps.v_out})
            dev_free(p);
    }
    // 3. capture the same run, its buffers bumped out of one workspace
    RG_HIP(hipMalloc(&G->arena, G->arena_bytes));
    RG_HIP(hipStreamCreateWithFlags(&G->cs, hipStreamNonBlocking));
    RG_HIP(hipStreamBeginCapture(G->cs, hipStreamCaptureModeThreadLocal));
    arena_begin(G->arena, G->arena_bytes);
    const int st = build_batched(pairs, npairs, &t, nullptr, G->cs, &G->caps);
    arena_end();
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(G->cs, &graph);
    G->graph = graph;
    if (st != RG_OK) return st;
    RG_HIP(e);
    RG_HIP(hipGraphInstantiate(&G->exec, G->graph, nullptr, nullptr, 0));
    *out = G.release();
    return RG_OK;
}

}  // namespace rg

rg_arena::~rg_arena()
{
    for (void* p : bufs) rg::dev_free(p);
}

namespace rg {

void weights_release(rg_weights* W)
{
    if (!W) return;
    // applies still queued on streams other than the handle's own read its
    // arrays: the memory returns to its own stream's pool only after them
    for (auto& u : W->uses) {
        (void)hipEventSynchronize(u.second);
        (void)hipEventDestroy(u.second);
    }
    if (!W->arena) {
        dev_free(W->rowptr_in);
        dev_free(W->col_in);
        dev_free(W->val_in);
        dev_free(W->rowptr_out);
        dev_free(W->col_out);
        dev_free(W->val_out);
    }
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
    W->home = s;
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

int rg_build_conservative_2d_batched(const rg_pair_desc* pairs, int64_t npairs,
                                     const rg_tuning* tuning, rg_weights_t* out,
                                     void* stream)
{
    return build_batched(pairs, npairs, tuning, out, (hipStream_t)stream);
}

int rg_tune_build(const rg_pair_desc* pairs, int64_t npairs,
                  const int32_t* edges_per_lane, int n_epl,
                  const int32_t* block, int n_block, int reps,
                  const rg_tuning* base, rg_tuning* best, double* best_ms,
                  double* times, void* stream)
{
    if (!edges_per_lane || !block || n_epl <= 0 || n_block <= 0 || !best)
        return fail(RG_ERR_ARGUMENT, "empty candidate list or NULL output");
    if (reps < 1) reps = 1;
    if (reps > 64) reps = 64;
    hipStream_t s = (hipStream_t)stream;
    std::vector<rg_weights*> hs((size_t)npairs > 0 ? (size_t)npairs : 1, nullptr);
    hipEvent_t e0, e1;
    RG_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(RG_ERR_HIP, "hipEventCreate failed");
    }
    double top = -1;
    int st = RG_OK;
    for (int a = 0; a < n_epl && st == RG_OK; a++) {
        for (int b = 0; b < n_block && st == RG_OK; b++) {
            rg_tuning t = base ? *base : rg_tuning{};
            t.edges_per_lane = edges_per_lane[a];
            t.block = block[b];
            t.flags &= ~RG_TUNE_TIMINGS;
            std::vector<double> ms;
            for (int r = 0; r <= reps && st == RG_OK; r++) {
                if (hipEventRecord(e0, s) != hipSuccess) st = fail(RG_ERR_HIP, "event record");
                if (st == RG_OK) st = build_batched(pairs, npairs, &t, hs.data(), s);
                if (st == RG_OK && hipEventRecord(e1, s) != hipSuccess)
                    st = fail(RG_ERR_HIP, "event record");
                for (auto& h : hs) {
                    weights_release(h);
                    h = nullptr;
                }
                if (st != RG_OK) break;
                float f = 0;
                if (hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&f, e0, e1) != hipSuccess) {
                    st = fail(RG_ERR_HIP, "event timing");
                    break;
                }
                if (r > 0) ms.push_back(f);  // the first run is untimed
            }
            if (st != RG_OK) break;
            std::sort(ms.begin(), ms.end());
            const double med = ms[ms.size() / 2];
            if (times) times[a * n_block + b] = med;
            if (top < 0 || med < top) {
                top = med;
                *best = t;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st != RG_OK) return st;
    if (best_ms) *best_ms = top;
    return RG_OK;
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

int rg_graph_create(const rg_pair_desc* pairs, int64_t npairs, const rg_tuning* tuning,
                    const rg_graph_job* jobs, double slack, rg_graph_t* out, void* stream)
{
    if (!out) return fail(RG_ERR_ARGUMENT, "out is NULL");
    static_assert(sizeof(rg_graph_job) == sizeof(GraphJob), "job layout");
    GraphImpl* G = nullptr;
    RG_TRY(graph_create(pairs, npairs, tuning, (const GraphJob*)jobs, slack, &G,
                        (hipStream_t)stream));
    *out = (rg_graph_t)G;
    return RG_OK;
}

int rg_graph_launch(rg_graph_t g, void* stream)
{
    if (!g) return fail(RG_ERR_ARGUMENT, "graph is NULL");
    GraphImpl* G = (GraphImpl*)g;
    RG_HIP(hipGraphLaunch(G->exec, (hipStream_t)stream));
    G->launched = true;
    return RG_OK;
}

int rg_graph_status(rg_graph_t g, int32_t* flags, void* stream)
{
    if (!g || !flags) return fail(RG_ERR_ARGUMENT, "graph or flags is NULL");
    GraphImpl* G = (GraphImpl*)g;
    if (!G->launched) return fail(RG_ERR_ARGUMENT, "the graph has not been launched");
    return read_scalars(G->caps.counters + 8, flags, sizeof(int32_t), (hipStream_t)stream);
}

int rg_graph_weights(rg_graph_t g, int64_t k, rg_weights_t* out, void* stream)
{
    if (!g || !out) return fail(RG_ERR_ARGUMENT, "graph or out is NULL");
    GraphImpl* G = (GraphImpl*)g;
    if (k < 0 || k >= G->npairs) return fail(RG_ERR_ARGUMENT, "pair index out of range");
    return graph_weights_copy(G, k, out, (hipStream_t)stream);
}

int rg_graph_capacity(rg_graph_t g, int64_t* triples, int64_t* overflow, int64_t* weights,
                      int64_t* workspace_bytes)
{
    if (!g) return fail(RG_ERR_ARGUMENT, "graph is NULL");
    GraphImpl* G = (GraphImpl*)g;
    if (triples) *triples = G->caps.T_cap;
    if (overflow) *overflow = G->caps.ov_cap;
    if (weights) *weights = G->caps.nnz_cap;
    if (workspace_bytes) *workspace_bytes = (int64_t)G->arena_bytes;
    return RG_OK;
}

int rg_graph_destroy(rg_graph_t g)
{
    delete (GraphImpl*)g;
    return RG_OK;
}

}  // extern "C"
