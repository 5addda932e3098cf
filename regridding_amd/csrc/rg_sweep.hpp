// rg_sweep.hpp -- the Ramshaw edge sweep, decomposed into independent chunks.
//
// The reference walks each row of the sweep grid serially
// (_weights_conservative_2d.py:286-387).  At every sweep VERTEX the walk state
// collapses to "outside the static grid" or "inside static cell c", with the
// last-crossed-edge memory reset (the no-hit branches at :540 and :706-710),
// and the walk position is exactly the vertex.  A row can therefore be cut at
// any vertex: the piece that starts there only needs that collapsed state.
//
// The GPU build gives each lane a chunk of K consecutive edges of one row.
// Chunk 0 of every row starts from the reference's own row-start state
// (bounding box, boundary polygon and secant search, :308-322).  Later chunks
// start from a guess (the secant cell search at their first vertex).  After
// all chunks have walked, a per-row pass compares every chunk's start state
// with the end state of its predecessor; a chunk whose guess was wrong is
// walked again from the true state, so the result is exactly the serial walk
// whatever the guesses were.
//
// All functions are __host__ __device__: tests/emu runs them on the CPU
// against the oracle before any GPU launch.
#pragma once

#include "rg_geometry.hpp"

namespace rg {

// Boundary-edge blocking of the outside-step scan.  Every edge's bounding box
// lies inside its block's box, and every block's box inside its super-block's
// box, so skipping a box the segment's bounding box does not touch skips
// only edges for which the reference's first test (bounding boxes,
// geometry.py:422-433) already fails.
constexpr int BND_BLOCK = 16;
constexpr int BND_SUPER = 16;  // blocks per super-block

struct Box {
    double x0, y0, x1, y1;  // inclusive [x0, x1] x [y0, y1]
};

// A static (searched) grid with its derived tables, all in device memory.
struct StaticGrid {
    const double* x;
    const double* y;
    const double2* xy;       // the same vertices interleaved (x, y): one
                             // 16-byte load per cell corner in the walk
    int64_t nx, ny;          // vertices
    const double* bbox;      // 4 values: min x, min y, max x, max y (numpy
                             // semantics: NaN propagates)
    const double* bpx;       // boundary polygon, grid_boundary order
    const double* bpy;
    int64_t nbp;
    const double* be;        // boundary edges in scan order, 4 doubles each
    const int32_t* bcell;    // flat static cell of each boundary edge
    int64_t nbe;
    const Box* bblk;         // box per BND_BLOCK edges
    int64_t nblk;
    const Box* bsup;         // box per BND_SUPER blocks
    int64_t nsup;
    const Box* ball;         // box of all boundary edges
    const double2* pyr;      // y range (min, max) of the vertices of each
    int64_t npyr;            // run of 64 boundary-polygon edges
};

// One (pass, axis) sweep: rows of the aligned sweep grid walked across the
// static grid (_sweep_grid, :133-203).
struct Pass {
    const double* xs;        // sweep grid vertices, original orientation
    const double* ys;
    const double2* xys;      // the same vertices interleaved (one load each)
    int64_t sr, sk;          // element (row, k) of the aligned view
    int64_t nrows, nk;       // aligned view shape
    int64_t ncy_sweep;       // cells along the last axis of the sweep grid
    int32_t axis;            // 0: aligned view is the transpose
    int32_t sweep_input;     // 1: sweep = input grid, static = output grid
    StaticGrid st;
    int64_t chunk_base;      // global id of the first chunk of this pass
    int64_t cpr;             // chunks per row
    int64_t thread_base;     // first walk thread of this pass
};

struct Build {
    Pass pass[4];            // F axis 0, F axis 1, T axis 0, T axis 1
    int64_t K;               // edges per chunk
    int64_t nchunks;
    int64_t nrows_total;
    int64_t row_base[4];     // id of the first row of each pass
    int64_t nthreads;        // walk threads (a multiple of 64 per pass)
    // Where this grid pair sits in a batched build's concatenated index
    // spaces (all 0 for a single pair).  The ids above are pair-local.
    // Input and output cells are padded to multiples of 64 per pair, so
    // every group of 64 cells belongs to one pair.
    int64_t thread_gbase, chunk_gbase, row_gbase, cell_gbase, out_gbase;
};

// Walk threads and chunks.  Chunk ids are canonical -- pass, row, chunk
// along the row -- because emission ordinals follow them.  Threads are
// assigned differently: a wavefront takes chunk c of 64 ADJACENT rows, so
// its lanes walk neighbouring strips and touch neighbouring static cells
// (one row's consecutive chunks would spread a wave along the whole row).
// Thread t of pass p: block of rows rb = t / (cpr * 64), chunk
// c = (t / 64) % cpr, row = rb * 64 + t % 64 (idle past the last row).
RG_HD int64_t threads_of_pass(int64_t nrows, int64_t cpr)
{
    return (nrows + 63) / 64 * cpr * 64;
}

struct ChunkRef {
    int p;
    int64_t row, c, g;
    bool valid;
};

RG_HD ChunkRef chunk_of_thread(const Build& B, int64_t t)
{
    ChunkRef r;
    int p = 0;
    if (t >= B.pass[1].thread_base) p = 1;
    if (t >= B.pass[2].thread_base) p = 2;
    if (t >= B.pass[3].thread_base) p = 3;
    const Pass& P = B.pass[p];
    const int64_t u = t - P.thread_base;
    r.p = p;
    r.c = (u >> 6) % P.cpr;
    r.row = (u / (P.cpr * 64)) * 64 + (u & 63);
    r.valid = r.row < P.nrows;
    r.g = P.chunk_base + r.row * P.cpr + r.c;
    return r;
}

RG_HD int64_t thread_of_chunk(const Pass& P, int64_t row, int64_t c)
{
    return P.thread_base + ((row >> 6) * P.cpr + c) * 64 + (row & 63);
}

// Per-step budget of a chunk: a non-degenerate walk moves at most one cell
// per step and crosses fewer than 2 (nx + ny) cells per edge; the reference
// walk can cycle forever only on exactly degenerate corners.  A chunk that
// exceeds the budget fails the build instead of hanging the GPU.
RG_HD int64_t step_limit(const StaticGrid& s, int64_t nedges)
{
    return nedges * (2 * (s.nx + s.ny) + 8) + 64;
}

// Area of cell (a, b) of an (nx, ny) vertex grid.  The reference starts
// from 0 and adds, for the axis-0 sweep (x and y swapped on the transposed
// view), -T(a, b) then +T(a, b+1), then for the axis-1 sweep -S(a, b) then
// +S(a+1, b) (_grids.py:84-140); this is that sequence for one cell.
RG_HD double cell_area(const double* __restrict__ x,
                       const double* __restrict__ y, int64_t ny, int64_t a,
                       int64_t b)
{
    const int64_t v00 = a * ny + b, v10 = v00 + ny;
    double acc = 0.0;
    acc -= area_tri(y[v00], x[v00], y[v10], x[v10]);
    acc += area_tri(y[v00 + 1], x[v00 + 1], y[v10 + 1], x[v10 + 1]);
    acc -= area_tri(x[v00], y[v00], x[v00 + 1], y[v00 + 1]);
    acc += area_tri(x[v10], y[v10], x[v10 + 1], y[v10 + 1]);
    return acc;
}

// cell_area on the interleaved vertex copy (the same operations in the
// same order, one load per vertex).
RG_HD double cell_area_xy(const double2* __restrict__ xy, int64_t ny, int64_t a, int64_t b)
{
    const int64_t v00 = a * ny + b, v10 = v00 + ny;
    const double2 p00 = xy[v00], p10 = xy[v10], p01 = xy[v00 + 1], p11 = xy[v10 + 1];
    double acc = 0.0;
    acc -= area_tri(p00.y, p00.x, p10.y, p10.x);
    acc += area_tri(p01.y, p01.x, p11.y, p11.x);
    acc -= area_tri(p00.x, p00.y, p01.x, p01.y);
    acc += area_tri(p10.x, p10.y, p11.x, p11.y);
    return acc;
}

RG_HD int pass_of_chunk(const Build& B, int64_t g)
{
    int p = 0;
    if (g >= B.pass[1].chunk_base) p = 1;
    if (g >= B.pass[2].chunk_base) p = 2;
    if (g >= B.pass[3].chunk_base) p = 3;
    return p;
}

RG_HD int pass_of_row(const Build& B, int64_t r)
{
    int p = 0;
    if (r >= B.row_base[1]) p = 1;
    if (r >= B.row_base[2]) p = 2;
    if (r >= B.row_base[3]) p = 3;
    return p;
}

RG_HD double sweep_x(const Pass& P, int64_t row, int64_t k)
{
    return gload(P.xs, row * P.sr + k * P.sk);
}

RG_HD double sweep_y(const Pass& P, int64_t row, int64_t k)
{
    return gload(P.ys, row * P.sr + k * P.sk);
}

// Scan-order code of boundary edge e: 2 * face + axis, the value the
// reference stores in index_edge_last after entering (:532-535).
RG_HD int boundary_code(const StaticGrid& s, int64_t e)
{
    int64_t na = s.ny - 1;
    int64_t nb = s.nx - 1;
    if (e < na) return 0;          // axis 0, face 0 (i = 0)
    if (e < 2 * na) return 2;      // axis 0, face 1 (i = last)
    if (e < 2 * na + nb) return 1; // axis 1, face 0 (j = 0)
    return 3;                      // axis 1, face 1 (j = last)
}

// Boundary edge e of the outside scan (:462-537): its end points (q1, q2) as
// the reference orders them and its static cell.  Used to fill the tables.
RG_HD void boundary_edge(const double* x, const double* y, int64_t nx,
                         int64_t ny, int64_t e, double* q, int32_t* cell)
{
    int64_t na = ny - 1, nb = nx - 1;
    int64_t ncy = ny - 1;
    int64_t v1, v2, c;
    if (e < na) {                   // vertices (0, i0) -> (0, i0 + 1)
        int64_t i0 = e;
        v1 = i0; v2 = i0 + 1; c = i0;
    } else if (e < 2 * na) {        // (nx-1, i0) -> (nx-1, i0 + 1)
        int64_t i0 = e - na;
        v1 = (nx - 1) * ny + i0; v2 = v1 + 1; c = (nx - 2) * ncy + i0;
    } else if (e < 2 * na + nb) {   // (i0, 0) -> (i0 + 1, 0)
        int64_t i0 = e - 2 * na;
        v1 = i0 * ny; v2 = (i0 + 1) * ny; c = i0 * ncy;
    } else {                        // (i0, ny-1) -> (i0 + 1, ny-1)
        int64_t i0 = e - 2 * na - nb;
        v1 = i0 * ny + ny - 1; v2 = (i0 + 1) * ny + ny - 1; c = i0 * ncy + ncy - 1;
    }
    q[0] = x[v1]; q[1] = y[v1]; q[2] = x[v2]; q[3] = y[v2];
    *cell = (int32_t)c;
}

// Vertex k of the boundary polygon (_grids.py:167-215 order).
RG_HD int64_t boundary_vertex(int64_t nx, int64_t ny, int64_t k)
{
    int64_t a = nx - 1, b = ny - 1;
    if (k < a) return k * ny;                          // j = 0, i up
    k -= a;
    if (k < b) return (nx - 1) * ny + k;               // i = last, j up
    k -= b;
    if (k < a) return (nx - 1 - k) * ny + (ny - 1);    // j = last, i down
    k -= a;
    return (ny - 1 - k);                               // i = 0, j down
}

RG_HD bool box_touch(const Box& b, double xl, double yl, double xh, double yh)
{
    return xl <= b.x1 && b.x0 <= xh && yl <= b.y1 && b.y0 <= yh;
}

// _step_outside_static (:401-552): the earliest boundary crossing of
// p1 -> p2, ties resolved by scan order, ignoring the cell just exited.
// Returns true and (cell, code, t) on a crossing.
RG_HD bool outside_scan(const StaticGrid& s, double x1, double y1, double x2,
                        double y2, int32_t skip_cell, int32_t* cell, int* code,
                        double* t_hit)
{
    // 32-bit indices (boundary edges of a grid below 2^31 vertices)
    const double xl = x1 > x2 ? x2 : x1, xh = x1 > x2 ? x1 : x2;
    const double yl = y1 > y2 ? y2 : y1, yh = y1 > y2 ? y1 : y2;
    const int32_t nsup = (int32_t)s.nsup, nblk = (int32_t)s.nblk, nbe = (int32_t)s.nbe;
    const Box* const bsup = s.bsup;
    const Box* const bblk = s.bblk;
    const double2* const be = (const double2*)s.be;  // 2 per edge
    const int32_t* const bcell = s.bcell;
    double t_min = __builtin_huge_val();
    int32_t best = -1;
    if (!box_touch(gload(s.ball, 0), xl, yl, xh, yh)) return false;
    for (int32_t su = 0; su < nsup; su++) {
        if (!box_touch(gload(bsup, su), xl, yl, xh, yh)) continue;
        const int32_t b1 = min((su + 1) * BND_SUPER, nblk);
        for (int32_t bl = su * BND_SUPER; bl < b1; bl++) {
            if (!box_touch(gload(bblk, bl), xl, yl, xh, yh)) continue;
            const int32_t e1 = min((bl + 1) * BND_BLOCK, nbe);
            for (int32_t e = bl * BND_BLOCK; e < e1; e++) {
                const double2 qa = gload(be, 2 * e), qb = gload(be, 2 * e + 1);
                TU r = seg_params(x1, y1, x2, y2, qa.x, qa.y, qb.x, qb.y);
                if (seg_hit(r) && r.t < t_min) {
                    if (gload(bcell, e) == skip_cell) continue;
                    t_min = r.t;
                    best = e;
                }
            }
        }
    }
    if (best >= 0) {
        *cell = gload(bcell, best);
        *code = boundary_code(s, best);
        *t_hit = t_min;
    }
    return best >= 0;
}

// A staged triple.  The count walk keeps up to `cap` triples per walk
// thread, slot-major within each wavefront's region: slot n of the 64 lanes
// of a wavefront is one contiguous 1 KiB run, so lanes emitting together
// write (and the scatter later reads) whole cache lines.  A lane-contiguous
// layout measured 1.6x the write and 2.3x the read traffic of this one.
// `a` is the signed line-integral term of
// the triple -- the weight before the division by the input cell's area,
// which the reduction applies (one area per bucket instead of one random
// load per triple).
struct alignas(16) Triple {
    int32_t cin, cout;
    double a;
};

RG_HD int64_t stage_index(int64_t t, int64_t n, int64_t cap)
{
    return ((t >> 6) * cap + n) * 64 + (t & 63);
}

// A bucket record: key (out << 32 | emission ordinal) and value.
struct alignas(16) KV {
    uint64_t key;
    double val;
};

// Walk end/start states: -1 outside, otherwise the flat static cell.
constexpr int32_t STATE_OUTSIDE = -1;
constexpr int32_t STATE_FAILED = -2;   // step budget exceeded

// Walk edges [k0, k1) of `row` from `start`, calling sink.emit(in, out, a)
// for every triple in the reference's order (left then right per
// sub-segment, :786-830), `a` being the signed line-integral term whose
// quotient by the input cell's area is the weight.  Returns the state at
// vertex k1.
template <class Sink>
RG_HD int32_t walk_chunk(const Build& B, const Pass& P, int64_t row64,
                         int64_t k0, int64_t k1, int32_t start, Sink& sink)
{
    // Indices are 32-bit: a build requires fewer than 2^31 vertices per grid
    // (so every vertex, cell and sweep index below fits).
    // Everything the steps read from the pass is copied to locals first: the
    // stores and atomics of the sink could alias the pass table as far as the
    // compiler knows, and it would otherwise reload those fields (and wait
    // for them) in every step.
    const StaticGrid& s = P.st;
    const double2* const sxy = s.xy;
    const double2* const wxy = P.xys;
    const bool axis0 = P.axis == 0;
    const bool sweep_in = P.sweep_input != 0;
    const int32_t ncx = (int32_t)(s.nx - 1), ncy = (int32_t)(s.ny - 1), nys = (int32_t)s.ny;
    const int32_t row = (int32_t)row64, kend = (int32_t)k1;
    const int32_t ncy_sweep = (int32_t)P.ncy_sweep, nrows = (int32_t)P.nrows;
    const int32_t sk = (int32_t)P.sk;
    // steps left in the budget (one register, counted down)
    int32_t budget = (int32_t)min(step_limit(s, k1 - k0), (int64_t)INT32_MAX - 1);
    int32_t k = (int32_t)k0;
    int32_t widx = row * (int32_t)P.sr + k * sk;  // sweep vertex (row, k)
    const double2 sv0 = gload(wxy, widx);
    double x1 = sv0.x, y1 = sv0.y;
    // the end of the current sweep edge, loaded when the walk reaches a
    // new edge (a crossing keeps it).  (Requesting the next edge's end one
    // step ahead costs registers -- a wavefront per SIMD -- and measured
    // slower.)
    double2 ev = make_double2(0, 0);
    if (k < kend) ev = gload(wxy, widx + sk);
    bool outside = start < 0;
    int32_t sx = outside ? 0 : start / ncy;
    int32_t sy = outside ? 0 : start % ncy;
    int32_t edge_last = -1;  // sys.maxsize
    // How the walk reached the current cell, i.e. which cached corners still
    // hold: the edge crossed out of the previous cell (0..3: the previous
    // cell's corners shift over and two are loaded), -1 (same cell: all
    // four hold) or -2 (none: all four are loaded).
    int32_t mv = -2;
    double2 c0, c1, c2, c3;
    c0.x = c0.y = c1.x = c1.y = c2.x = c2.y = c3.x = c3.y = 0;
    while (k < kend) {
        if (--budget < 0) return STATE_FAILED;
        double x2 = ev.x, y2 = ev.y;
        if (outside) {
            int32_t cell;
            int code;
            double t;
            // the cached corners are reloaded after an outside stretch
            mv = -2;
            c0.x = c0.y = c1.x = c1.y = c2.x = c2.y = c3.x = c3.y = 0;
            if (outside_scan(s, x1, y1, x2, y2, edge_last, &cell, &code, &t)) {
                x2 = x1 + t * (x2 - x1);
                y2 = y1 + t * (y2 - y1);
                sx = cell / ncy;
                sy = cell % ncy;
                edge_last = code;
                outside = false;
            } else {
                k += 1;
                widx += sk;
                edge_last = -1;
                if (k < kend) ev = gload(wxy, widx + sk);
            }
        } else {
            // _step_inside_static (:561-749).  Cell corners c0..c3 =
            // (0,0) (1,0) (1,1) (0,1); edge v runs c[v-1] -> c[v] and its
            // outward neighbour is cell_normals[v].  The corners stay in
            // registers while the walk stays in the cell; a move to a
            // neighbour loads only the two new corners.
            if (mv != -1) {
                const int32_t a = sx * nys + sy, b = a + nys;
                if (mv == 2) {         // moved +x
                    c0 = c1; c3 = c2; c1 = gload(sxy, b); c2 = gload(sxy, b + 1);
                } else if (mv == 0) {  // -x
                    c1 = c0; c2 = c3; c0 = gload(sxy, a); c3 = gload(sxy, a + 1);
                } else if (mv == 3) {  // +y
                    c0 = c3; c1 = c2; c3 = gload(sxy, a + 1); c2 = gload(sxy, b + 1);
                } else if (mv == 1) {  // -y
                    c3 = c0; c2 = c1; c0 = gload(sxy, a); c1 = gload(sxy, b);
                } else {
                    c0 = gload(sxy, a); c1 = gload(sxy, b); c2 = gload(sxy, b + 1);
                    c3 = gload(sxy, a + 1);
                }
            }
            const double cx0 = c0.x, cy0 = c0.y, cx1 = c1.x, cy1 = c1.y;
            const double cx2 = c2.x, cy2 = c2.y, cx3 = c3.x, cy3 = c3.y;
            int hit = -1;
            TU r;
            r.t = 0;
            if (edge_last != 0) {
                r = seg_params(x1, y1, x2, y2, cx3, cy3, cx0, cy0);
                if (seg_hit(r)) hit = 0;
            }
            if (hit < 0 && edge_last != 1) {
                r = seg_params(x1, y1, x2, y2, cx0, cy0, cx1, cy1);
                if (seg_hit(r)) hit = 1;
            }
            if (hit < 0 && edge_last != 2) {
                r = seg_params(x1, y1, x2, y2, cx1, cy1, cx2, cy2);
                if (seg_hit(r)) hit = 2;
            }
            if (hit < 0 && edge_last != 3) {
                r = seg_params(x1, y1, x2, y2, cx2, cy2, cx3, cy3);
                if (seg_hit(r)) hit = 3;
            }
            int32_t nsx = sx, nsy = sy, k_next = k;
            int32_t e_next;
            if (hit >= 0) {
                x2 = x1 + r.t * (x2 - x1);
                y2 = y1 + r.t * (y2 - y1);
                nsx += (hit == 0) ? -1 : (hit == 2 ? 1 : 0);
                nsy += (hit == 1) ? -1 : (hit == 3 ? 1 : 0);
                e_next = (hit + 2) & 3;
            } else {
                k_next = k + 1;
                e_next = -1;
            }
            // _calc_and_save_weights (:757-830) for p1 -> p2 at sweep
            // index (row, k) and static cell (sx, sy).
            double area = area_tri(x1, y1, x2, y2);
            if (axis0) area = -area;
            const int32_t cell_static = sx * ncy + sy;
            for (int side = 0; side < 2; side++) {
                const int32_t rr = side == 0 ? row - 1 : row;
                if (side == 0 && !(rr >= 0)) continue;
                if (side == 1 && !(rr < nrows - 1)) continue;
                const int32_t cell_sweep = axis0 ? k * ncy_sweep + rr : rr * ncy_sweep + k;
                const int32_t cin = sweep_in ? cell_sweep : cell_static;
                const int32_t cout = sweep_in ? cell_static : cell_sweep;
                // the weight is term / area(cin) [* weights_input[cin]]
                // (:797-802, :820-825); the reduction divides
                sink.emit(cin, cout, side == 0 ? area : -area);
            }
            if (!(nsx >= 0 && nsy >= 0 && nsx < ncx && nsy < ncy)) {
                e_next = cell_static;
                outside = true;
            } else {
                sx = nsx;
                sy = nsy;
            }
            mv = hit;  // (-1: still in the same cell)
            if (k_next != k) {
                widx += sk;
                if (k_next < kend) ev = gload(wxy, widx + sk);
            }
            k = k_next;
            edge_last = e_next;
        }
        x1 = x2;
        y1 = y2;
    }
    return outside ? STATE_OUTSIDE : sx * ncy + sy;
}

// Guessed walk state at the first vertex of chunk c > 0 of a row: the
// reference's own cell search (bounding box, then secant iteration), with
// "outside" for anything it does not place in a cell.  Only a guess: the
// per-row consistency pass corrects it.
RG_HD int32_t chunk_start_guess(const Pass& P, int64_t row, int64_t k)
{
    const StaticGrid& s = P.st;
    double px = sweep_x(P, row, k), py = sweep_y(P, row, k);
    const double bx0 = gload(s.bbox, 0), by0 = gload(s.bbox, 1);
    const double bx1 = gload(s.bbox, 2), by1 = gload(s.bbox, 3);
    if (!(bx0 <= px && px <= bx1 && by0 <= py && py <= by1))
        return STATE_OUTSIDE;
    int64_t c = index_secant_xy(s.xy, s.nx, s.ny, px, py);
    return c >= 0 ? (int32_t)c : STATE_OUTSIDE;
}

// The weight of a triple from its term: a / area(in), times
// weights_input[in] when given -- the operations of :797-802 / :820-825 in
// their order.
RG_HD double weight_of(double a, double vol, const double* wi, int64_t cin)
{
    double w = a / vol;
    if (wi) w *= wi[cin];
    return w;
}

// numpy's pairwise summation of a[0..n) (the float64 add-reduce inner loop
// used by np.add.reduceat): under 8 terms a sequential sum seeded with -0.0,
// up to 128 terms eight interleaved accumulators, above that a split at n/2
// rounded down to a multiple of 8.  The split recursion is evaluated with an
// explicit stack (depth <= 64 for any int64 n).
// (`st` is the element stride of `a`, 2 for the values of KV records.)
RG_HD double pairwise_leaf(const double* a, int64_t n, int64_t st = 1)
{
    if (n < 8) {
        double res = -0.0;
        for (int64_t i = 0; i < n; i++) res += a[i * st];
        return res;
    }
    double r0 = a[0], r1 = a[st], r2 = a[2 * st], r3 = a[3 * st];
    double r4 = a[4 * st], r5 = a[5 * st], r6 = a[6 * st], r7 = a[7 * st];
    int64_t i;
    for (i = 8; i < n - (n % 8); i += 8) {
        r0 += a[(i + 0) * st]; r1 += a[(i + 1) * st];
        r2 += a[(i + 2) * st]; r3 += a[(i + 3) * st];
        r4 += a[(i + 4) * st]; r5 += a[(i + 5) * st];
        r6 += a[(i + 6) * st]; r7 += a[(i + 7) * st];
    }
    double res = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
    for (; i < n; i++) res += a[i * st];
    return res;
}

RG_HD double pairwise_sum(const double* a, int64_t n, int64_t stride = 1)
{
    if (n <= 128) return pairwise_leaf(a, n, stride);
    // post-order traversal of the split tree
    struct Frame {
        int64_t off, n;
        int stage;
        double left;
    };
    Frame fs[64];
    int top = 0;
    fs[0].off = 0; fs[0].n = n; fs[0].stage = 0; fs[0].left = 0;
    double ret = 0;
    while (top >= 0) {
        Frame& f = fs[top];
        if (f.n <= 128) {
            ret = pairwise_leaf(a + f.off * stride, f.n, stride);
            top--;
            continue;
        }
        int64_t n2 = f.n / 2;
        n2 -= n2 % 8;
        if (f.stage == 0) {
            f.stage = 1;
            Frame& c = fs[top + 1];
            c.off = f.off; c.n = n2; c.stage = 0; c.left = 0;
            top++;
        } else if (f.stage == 1) {
            f.left = ret;
            f.stage = 2;
            Frame& c = fs[top + 1];
            c.off = f.off + n2; c.n = f.n - n2; c.stage = 0; c.left = 0;
            top++;
        } else {
            ret = f.left + ret;
            top--;
        }
    }
    return ret;
}

// numpy's split of a pairwise-sum node of n > 128 elements: the left child
// takes n/2 rounded down to a multiple of 8.
RG_HD int64_t pairwise_split(int64_t n)
{
    const int64_t n2 = n / 2;
    return n2 - n2 % 8;
}

// The same sum as pairwise_sum, without a stack (no private memory on the
// GPU): the split tree is walked in post-order, each finished subtree's sum
// is stored in its first element, and the parent of a node is found again by
// descending from the root.  Destroys a[0 .. n) (strided).
RG_HD double pairwise_sum_inplace(double* a, int64_t n, int64_t stride = 1)
{
    if (n <= 128) return pairwise_leaf(a, n, stride);
    int64_t off = 0, len = n;
    while (len > 128) len = pairwise_split(len);  // leftmost leaf
    for (;;) {
        a[off * stride] = pairwise_leaf(a + off * stride, len, stride);
        for (;;) {
            if (len == n) return a[0];
            // parent (po, pl) of node (off, len), and which child it is
            int64_t po = 0, pl = n;
            bool left;
            for (;;) {
                const int64_t n2 = pairwise_split(pl);
                if (off == po && len == n2) {
                    left = true;
                    break;
                }
                if (off == po + n2 && len == pl - n2) {
                    left = false;
                    break;
                }
                if (off < po + n2) {
                    pl = n2;
                } else {
                    po += n2;
                    pl -= n2;
                }
            }
            if (left) {
                // on to the right sibling's leftmost leaf
                const int64_t n2 = pairwise_split(pl);
                off = po + n2;
                len = pl - n2;
                while (len > 128) len = pairwise_split(len);
                break;
            }
            // both children done: left sum at po, right sum at off
            a[po * stride] = a[po * stride] + a[off * stride];
            off = po;
            len = pl;
        }
    }
}

// np.add.reduceat over one segment: first element plus the pairwise sum of
// the rest (_weights_arrays.py:72).
RG_HD double reduceat_segment(const double* a, int64_t n, int64_t st = 1)
{
    if (n == 1) return a[0];
    return a[0] + pairwise_sum(a + st, n - 1, st);
}

// Sort keys[0..n) ascending with vals following.  Keys are unique, so the
// result is fully determined.  Insertion sort for short ranges, heapsort
// beyond (bounded work, no extra memory).
template <class K, class V>
RG_HD void sort_pairs(K* keys, V* vals, int64_t n)
{
    if (n <= 48) {
        for (int64_t i = 1; i < n; i++) {
            K k = keys[i];
            V v = vals[i];
            int64_t j = i - 1;
            while (j >= 0 && keys[j] > k) {
                keys[j + 1] = keys[j];
                vals[j + 1] = vals[j];
                j--;
            }
            keys[j + 1] = k;
            vals[j + 1] = v;
        }
        return;
    }
    // heapsort
    for (int64_t start = n / 2 - 1; start >= 0; start--) {
        int64_t root = start;
        while (2 * root + 1 < n) {
            int64_t child = 2 * root + 1;
            if (child + 1 < n && keys[child] < keys[child + 1]) child++;
            if (keys[root] < keys[child]) {
                K tk = keys[root]; keys[root] = keys[child]; keys[child] = tk;
                V tv = vals[root]; vals[root] = vals[child]; vals[child] = tv;
                root = child;
            } else {
                break;
            }
        }
    }
    for (int64_t end = n - 1; end > 0; end--) {
        K tk = keys[0]; keys[0] = keys[end]; keys[end] = tk;
        V tv = vals[0]; vals[0] = vals[end]; vals[end] = tv;
        int64_t root = 0;
        while (2 * root + 1 < end) {
            int64_t child = 2 * root + 1;
            if (child + 1 < end && keys[child] < keys[child + 1]) child++;
            if (keys[root] < keys[child]) {
                K tk2 = keys[root]; keys[root] = keys[child]; keys[child] = tk2;
                V tv2 = vals[root]; vals[root] = vals[child]; vals[child] = tv2;
                root = child;
            } else {
                break;
            }
        }
    }
}

// Sort bucket records by key (unique): insertion sort for short ranges,
// heapsort beyond.
RG_HD void sort_records(KV* r, int64_t n)
{
    if (n <= 48) {
        for (int64_t i = 1; i < n; i++) {
            KV x = r[i];
            int64_t j = i - 1;
            while (j >= 0 && r[j].key > x.key) {
                r[j + 1] = r[j];
                j--;
            }
            r[j + 1] = x;
        }
        return;
    }
    for (int64_t start = n / 2 - 1; start >= 0; start--) {
        int64_t root = start;
        while (2 * root + 1 < n) {
            int64_t child = 2 * root + 1;
            if (child + 1 < n && r[child].key < r[child + 1].key) child++;
            if (r[root].key < r[child].key) {
                KV t = r[root]; r[root] = r[child]; r[child] = t;
                root = child;
            } else {
                break;
            }
        }
    }
    for (int64_t end = n - 1; end > 0; end--) {
        KV t = r[0]; r[0] = r[end]; r[end] = t;
        int64_t root = 0;
        while (2 * root + 1 < end) {
            int64_t child = 2 * root + 1;
            if (child + 1 < end && r[child].key < r[child + 1].key) child++;
            if (r[root].key < r[child].key) {
                KV u = r[root]; r[root] = r[child]; r[child] = u;
                root = child;
            } else {
                break;
            }
        }
    }
}

// Coalesce one input cell's bucket: keys are (out << 32 | ordinal), values
// the emitted weights.  After sorting, runs of equal `out` are the duplicate
// (in, out) pairs in emission order; each run is reduced in reduceat order
// (_weights_arrays.py:59-72).  The unique (out, weight) pairs are written
// back to the front of the bucket; returns their count.
//
// A run of more than 129 terms needs the recursive split of the pairwise sum
// (with a stack in scratch memory).  Such runs do not occur for cells that
// overlap in a polygon of a few edges; if one does, the bucket is left
// sorted but not compacted, *long_run is set, and bucket_reduce_long (run by
// a separate kernel) finishes it.
// (entry j of the bucket at keys[j * st], vals[j * st])
RG_HD int64_t bucket_reduce_sorted(uint64_t* keys, double* vals, int64_t n,
                                   bool* long_run, int64_t st = 1)
{
    // Both passes stream over the terms and load them four at a time: a scan
    // that loads the next key only after comparing the previous one waits out
    // one memory (LDS) latency per term.  First the run lengths: a run of more
    // than 129 terms leaves the cell sorted and unreduced for the long-run
    // pass, and then the number of runs is returned.
    int64_t run = 0, longest = 0, m = 0;
    uint64_t prev = 0;
    for (int64_t t0 = 0; t0 < n; t0 += 4) {
        uint64_t o[4];
#pragma unroll
        for (int u = 0; u < 4; u++) o[u] = t0 + u < n ? keys[(t0 + u) * st] >> 32 : 0;
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (t0 + u < n) {
                const bool same = t0 + u > 0 && o[u] == prev;
                run = same ? run + 1 : 0;
                m += same ? 0 : 1;
                longest = run > longest ? run : longest;
                prev = o[u];
            }
    }
    *long_run = longest > 128;
    if (*long_run) return m;
    // Then each run becomes its first term plus the sum of the others.  Under
    // 8 others that sum is numpy's sequential one seeded with -0.0 and is
    // accumulated as the terms stream by; a longer run sums its terms with
    // pairwise_leaf, still in place (a run's output goes to a slot at or
    // before its first term, after its terms are read).
    m = 0;
    int64_t p = 0, cnt = 0;
    uint64_t cur = 0;
    double first = 0, res = -0.0;
    for (int64_t t0 = 0; t0 < n; t0 += 4) {
        uint64_t o[4];
        double v[4];
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (t0 + u < n) {
                o[u] = keys[(t0 + u) * st] >> 32;
                v[u] = vals[(t0 + u) * st];
            }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int64_t t = t0 + u;
            if (t >= n) continue;
            if (t > 0 && o[u] == cur) {
                cnt++;
                res += v[u];
                continue;
            }
            if (t > 0) {
                keys[m * st] = cur;
                vals[m * st] = cnt == 0  ? first
                               : cnt < 8 ? first + res
                                         : first + pairwise_leaf(vals + (p + 1) * st, cnt, st);
                m++;
            }
            cur = o[u];
            first = v[u];
            res = -0.0;
            cnt = 0;
            p = t;
        }
    }
    if (n > 0) {
        keys[m * st] = cur;
        vals[m * st] = cnt == 0  ? first
                       : cnt < 8 ? first + res
                                 : first + pairwise_leaf(vals + (p + 1) * st, cnt, st);
        m++;
    }
    return m;
}

RG_HD int64_t bucket_reduce(uint64_t* keys, double* vals, int64_t n,
                            bool* long_run)
{
    sort_pairs(keys, vals, n);
    return bucket_reduce_sorted(keys, vals, n, long_run);
}

// Sorting network over N keys in registers (N a power of two, every index
// static once unrolled), values following.
template <int N, class K>
RG_HD void sort_network(K* k, double* v)
{
    // Batcher's odd-even merge sort (63 comparators for 16 keys where the
    // bitonic network takes 80, 19 for 8 where it takes 24), unrolled into
    // registers; the keys come out ascending
#pragma unroll
    for (int p = 1; p < N; p <<= 1) {
#pragma unroll
        for (int q = p; q > 0; q >>= 1) {
#pragma unroll
            for (int j = q % p; j + q < N; j += 2 * q) {
#pragma unroll
                for (int i = 0; i < q; i++) {
                    const int a = i + j, b = i + j + q;
                    if (b < N && a / (2 * p) == b / (2 * p) && k[a] > k[b]) {
                        const K tk = k[a];
                        k[a] = k[b];
                        k[b] = tk;
                        const double tv = v[a];
                        v[a] = v[b];
                        v[b] = tv;
                    }
                }
            }
        }
    }
}

// Completion of a bucket left sorted (KV records, weights already divided)
// by bucket_reduce: any run length; the values of each run are consumed in
// place.
RG_HD int64_t bucket_reduce_long(KV* r, int64_t n)
{
    int64_t m = 0, p = 0;
    while (p < n) {
        uint64_t out = r[p].key >> 32;
        int64_t q = p + 1;
        while (q < n && (r[q].key >> 32) == out) q++;
        // reduceat_segment in place: a0 + pairwise(a1..)
        double* a = &r[p].val;
        double v = q - p == 1 ? a[0] : a[0] + pairwise_sum_inplace(a + 2, q - p - 1, 2);
        r[m].key = out;
        r[m].val = v;
        m++;
        p = q;
    }
    return m;
}

}  // namespace rg
