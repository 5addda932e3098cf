// rg_emu.cpp -- CPU emulation of the GPU build (test infrastructure).
//
// Runs the per-lane bodies of the GPU kernels (regridding_amd/csrc/
// rg_sweep.hpp: row start, chunk start guess, chunked walk, consistency
// fix-up, bucket reduction) serially on the host, with the same chunk
// decomposition and order of phases as rg_build.hip.  The tests compare its
// emission sequence and coalesced weights with the oracle bit for bit, which
// checks the chunk decomposition and every device function without a GPU.
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../regridding_amd/csrc/rg_layout.hpp"
#include "../../regridding_amd/csrc/rg_sweep.hpp"

using namespace rg;

namespace {

struct HStage {
    int32_t n;
    int64_t g, cap;
    Triple* stage;
    int32_t* hist;
    int32_t p;
    __host__ __device__ void emit(int64_t cin, int64_t cout, double w)
    {
        if (n < cap) {
            Triple t;
            t.cin = (int32_t)cin;
            t.cout = (int32_t)cout;
            t.a = w;
            stage[stage_index(g, n, cap)] = t;
        }
        n++;
        hist[4 * cin + p]++;
    }
};

struct HUncount {
    int32_t* hist;
    int32_t p;
    __host__ __device__ void emit(int64_t cin, int64_t, double) { hist[4 * cin + p]--; }
};

struct HFlat {
    int64_t pos;
    int64_t* in;
    int64_t* out;
    double* w;
    const double* vol;
    const double* wi;
    __host__ __device__ void emit(int64_t cin, int64_t cout, double v)
    {
        in[pos] = cin;
        out[pos] = cout;
        w[pos] = weight_of(v, vol[cin], wi, cin);
        pos++;
    }
};

struct HBucket {
    uint32_t ordinal;
    const int64_t* off;
    int32_t* cursor;
    uint64_t* key;
    double* val;
    int32_t pass;
    __host__ __device__ void emit(int64_t cin, int64_t cout, double v)
    {
        int64_t p = off[4 * cin + pass] + cursor[4 * cin + pass]++;
        key[p] = ((uint64_t)cout << 32) | ordinal;
        val[p] = v;
        ordinal++;
    }
};

double nmin(double a, double b)
{
    return (a != a || b != b) ? __builtin_nan("") : (a < b ? a : b);
}
double nmax(double a, double b)
{
    return (a != a || b != b) ? __builtin_nan("") : (a > b ? a : b);
}

struct Tables {
    std::vector<double2> xy;
    std::vector<double> bbox, bpx, bpy, be;
    std::vector<int32_t> bcell;
    std::vector<Box> bblk, bsup, ball;
};

void prepare(const double* x, const double* y, int64_t nx, int64_t ny,
             Tables& T, StaticGrid& G)
{
    int64_t nv = nx * ny, nbe = 2 * (nx - 1) + 2 * (ny - 1);
    T.xy.resize(nv);
    for (int64_t i = 0; i < nv; i++) {
        T.xy[i].x = x[i];
        T.xy[i].y = y[i];
    }
    T.bbox = {x[0], y[0], x[0], y[0]};
    for (int64_t i = 0; i < nv; i++) {
        T.bbox[0] = nmin(T.bbox[0], x[i]);
        T.bbox[1] = nmin(T.bbox[1], y[i]);
        T.bbox[2] = nmax(T.bbox[2], x[i]);
        T.bbox[3] = nmax(T.bbox[3], y[i]);
    }
    T.bpx.resize(nbe);
    T.bpy.resize(nbe);
    T.be.resize(4 * nbe);
    T.bcell.resize(nbe);
    for (int64_t e = 0; e < nbe; e++) {
        int64_t v = boundary_vertex(nx, ny, e);
        T.bpx[e] = x[v];
        T.bpy[e] = y[v];
        boundary_edge(x, y, nx, ny, e, &T.be[4 * e], &T.bcell[e]);
    }
    int64_t nblk = (nbe + BND_BLOCK - 1) / BND_BLOCK;
    int64_t nsup = (nblk + BND_SUPER - 1) / BND_SUPER;
    T.bblk.resize(nblk);
    for (int64_t b = 0; b < nblk; b++) {
        Box r{__builtin_huge_val(), __builtin_huge_val(), -__builtin_huge_val(),
              -__builtin_huge_val()};
        for (int64_t e = b * BND_BLOCK; e < std::min(nbe, (b + 1) * BND_BLOCK); e++) {
            const double* q = &T.be[4 * e];
            r.x0 = fmin(r.x0, fmin(q[0], q[2]));
            r.x1 = fmax(r.x1, fmax(q[0], q[2]));
            r.y0 = fmin(r.y0, fmin(q[1], q[3]));
            r.y1 = fmax(r.y1, fmax(q[1], q[3]));
        }
        T.bblk[b] = r;
    }
    T.bsup.resize(nsup);
    for (int64_t s = 0; s < nsup; s++) {
        Box r{__builtin_huge_val(), __builtin_huge_val(), -__builtin_huge_val(),
              -__builtin_huge_val()};
        for (int64_t b = s * BND_SUPER; b < std::min(nblk, (s + 1) * BND_SUPER); b++) {
            r.x0 = fmin(r.x0, T.bblk[b].x0);
            r.y0 = fmin(r.y0, T.bblk[b].y0);
            r.x1 = fmax(r.x1, T.bblk[b].x1);
            r.y1 = fmax(r.y1, T.bblk[b].y1);
        }
        T.bsup[s] = r;
    }
    Box all{__builtin_huge_val(), __builtin_huge_val(), -__builtin_huge_val(),
            -__builtin_huge_val()};
    for (const Box& r : T.bsup) {
        all.x0 = fmin(all.x0, r.x0);
        all.y0 = fmin(all.y0, r.y0);
        all.x1 = fmax(all.x1, r.x1);
        all.y1 = fmax(all.y1, r.y1);
    }
    T.ball.assign(1, all);
    G.x = x; G.y = y; G.xy = T.xy.data(); G.nx = nx; G.ny = ny;
    G.bbox = T.bbox.data();
    G.bpx = T.bpx.data(); G.bpy = T.bpy.data(); G.nbp = nbe;
    G.be = T.be.data(); G.bcell = T.bcell.data(); G.nbe = nbe;
    G.bblk = T.bblk.data(); G.nblk = nblk;
    G.bsup = T.bsup.data(); G.nsup = nsup;
    G.ball = T.ball.data();
}

// k_row_start, one wavefront serialised: winding terms summed in lane order
// (exact in any order), secant, exhaustive fallback = first containing cell.
int32_t row_start(const Pass& P, int64_t row, int64_t* anomalies)
{
    const StaticGrid& s = P.st;
    double px = sweep_x(P, row, 0), py = sweep_y(P, row, 0);
    if (!(s.bbox[0] <= px && px <= s.bbox[2] && s.bbox[1] <= py &&
          py <= s.bbox[3]))
        return STATE_OUTSIDE;
    double w = 0;
    for (int64_t k = 0; k < s.nbp; k++) {
        int64_t km = k == 0 ? s.nbp - 1 : k - 1;
        w += winding_term(px, py, s.bpx[km], s.bpy[km], s.bpx[k], s.bpy[k]);
    }
    if (w == 0) return STATE_OUTSIDE;
    int64_t c = index_secant(s.x, s.y, s.nx, s.ny, px, py);
    if (c == RG_SEARCH_BRUTE) {
        c = -1;
        int64_t ncy = s.ny - 1, nc = (s.nx - 1) * ncy;
        for (int64_t q = 0; q < nc; q++)
            if (cell_contains(s.x, s.y, s.ny, q / ncy, q % ncy, px, py)) {
                c = q;
                break;
            }
    }
    if (c < 0) {
        (*anomalies)++;
        return STATE_OUTSIDE;
    }
    return (int32_t)c;
}

void chunk_range(const Build& B, const Pass& P, int64_t c, int64_t* k0,
                 int64_t* k1)
{
    *k0 = c * B.K;
    int64_t e = *k0 + B.K;
    *k1 = e < P.nk - 1 ? e : P.nk - 1;
    if (*k1 < *k0) *k1 = *k0;
}

}  // namespace

extern "C" {

// Emulated build.  Outputs (malloc'd, free with rge_free):
//   flat_*   uncoalesced triples in emission-ordinal order (n_flat)
//   csr_*    coalesced triples sorted by (in, out) (n_csr)
// stats: [n_chunks, n_fixups, n_anomalies, n_long_buckets, failed, restaged,
//         thread/chunk map errors]
// `guess` selects the chunk start guesses: 0 = the GPU's cell-search guess,
// 1 = always "outside" (every chunk of a row that starts inside must then be
// repaired by the fix-up pass), 2 = pseudo-random garbage states.
int64_t rge_build(const double* xi, const double* yi, int64_t nxi, int64_t nyi,
                  const double* xo, const double* yo, int64_t nxo, int64_t nyo,
                  const double* wi, int64_t K, int64_t stage_per_edge, int guess,
                  int64_t** flat_in,
                  int64_t** flat_out, double** flat_w, int64_t* n_flat,
                  int64_t** csr_in, int64_t** csr_out, double** csr_w,
                  int64_t* stats)
{
    const int64_t n_in = (nxi - 1) * (nyi - 1), n_out = (nxo - 1) * (nyo - 1);
    std::vector<double> vol(n_in);
    for (int64_t c = 0; c < n_in; c++)
        vol[c] = cell_area(xi, yi, nyi, c / (nyi - 1), c % (nyi - 1));
    Tables Ti, To;
    StaticGrid Gi, Go;
    prepare(xi, yi, nxi, nyi, Ti, Gi);
    prepare(xo, yo, nxo, nyo, To, Go);

    Build B;
    B.K = K;
    int64_t chunk = 0, rows = 0, threads = 0;
    for (int p = 0; p < 4; p++) {
        Pass& P = B.pass[p];
        const int sweep_input = p >= 2, axis = p & 1;
        const int64_t nxs = sweep_input ? nxi : nxo, nys = sweep_input ? nyi : nyo;
        P.xs = sweep_input ? xi : xo;
        P.ys = sweep_input ? yi : yo;
        P.xys = sweep_input ? Gi.xy : Go.xy;
        if (axis == 0) { P.sr = 1; P.sk = nys; P.nrows = nys; P.nk = nxs; }
        else { P.sr = nys; P.sk = 1; P.nrows = nxs; P.nk = nys; }
        P.ncy_sweep = nys - 1;
        P.axis = axis;
        P.sweep_input = sweep_input;
        P.st = sweep_input ? Go : Gi;
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
    // the thread <-> chunk maps: every chunk owned by exactly one thread
    int64_t map_errors = 0;
    {
        std::vector<int32_t> owner(chunk, 0);
        for (int64_t t = 0; t < threads; t++) {
            ChunkRef r = chunk_of_thread(B, t);
            if (!r.valid) continue;
            if (r.g < 0 || r.g >= chunk) { map_errors++; continue; }
            owner[r.g]++;
            if (thread_of_chunk(B.pass[r.p], r.row, r.c) != t) map_errors++;
        }
        for (int64_t g = 0; g < chunk; g++)
            if (owner[g] != 1) map_errors++;
    }

    std::vector<int32_t> start(chunk), end(chunk), count(chunk), hist(4 * n_in, 0);
    const int64_t cap = K * stage_per_edge;
    std::vector<Triple> stage((size_t)threads * cap);
    int64_t anomalies = 0, fixups = 0, failed = 0, nlong = 0, restaged = 0;
    uint64_t lcg = 0x9E3779B97F4A7C15ull;
    for (int64_t r = 0; r < rows; r++) {
        int p = pass_of_row(B, r);
        const Pass& P = B.pass[p];
        int64_t row = r - B.row_base[p];
        start[P.chunk_base + row * P.cpr] = row_start(P, row, &anomalies);
    }
    for (int64_t g = 0; g < chunk; g++) {
        int p = pass_of_chunk(B, g);
        const Pass& P = B.pass[p];
        int64_t local = g - P.chunk_base, row = local / P.cpr, c = local % P.cpr;
        if (c == 0) continue;
        if (guess == 0) {
            start[g] = chunk_start_guess(P, row, c * K);
        } else if (guess == 1) {
            start[g] = STATE_OUTSIDE;
        } else {
            lcg = lcg * 6364136223846793005ull + 1442695040888963407ull;
            int64_t ncells = (P.st.nx - 1) * (P.st.ny - 1);
            start[g] = (int32_t)((int64_t)(lcg >> 33) % (ncells + 1)) - 1;
        }
    }
    // count
    for (int64_t g = 0; g < chunk; g++) {
        int p = pass_of_chunk(B, g);
        const Pass& P = B.pass[p];
        int64_t local = g - P.chunk_base, row = local / P.cpr, c = local % P.cpr;
        int64_t k0, k1;
        chunk_range(B, P, c, &k0, &k1);
        HStage s{0, thread_of_chunk(P, row, c), cap, stage.data(), hist.data(), p};
        end[g] = walk_chunk(B, P, row, k0, k1, start[g], s);
        count[g] = s.n;
    }
    // fix-up (k_fixup)
    for (int64_t r = 0; r < rows; r++) {
        int p = pass_of_row(B, r);
        const Pass& P = B.pass[p];
        int64_t row = r - B.row_base[p], g0 = P.chunk_base + row * P.cpr;
        for (int64_t c = 1; c < P.cpr; c++) {
            int64_t g = g0 + c;
            int32_t truth = end[g - 1];
            if (truth == STATE_FAILED) break;
            if (start[g] == truth) continue;
            int64_t k0, k1;
            chunk_range(B, P, c, &k0, &k1);
            HUncount u{hist.data(), p};
            (void)walk_chunk(B, P, row, k0, k1, start[g], u);
            HStage rc{0, thread_of_chunk(P, row, c), cap, stage.data(), hist.data(), p};
            end[g] = walk_chunk(B, P, row, k0, k1, truth, rc);
            count[g] = rc.n;
            start[g] = truth;
            fixups++;
        }
        for (int64_t c = 0; c < P.cpr; c++)
            if (end[g0 + c] == STATE_FAILED) {
                failed++;
                break;
            }
    }
    std::vector<int64_t> choff(chunk + 1, 0), boff(4 * n_in + 1, 0);
    for (int64_t g = 0; g < chunk; g++) choff[g + 1] = choff[g] + count[g];
    for (int64_t i = 0; i < 4 * n_in; i++) boff[i + 1] = boff[i] + hist[i];
    const int64_t T = choff[chunk];
    // flat emission in ordinal order (for comparison with the oracle sweep)
    int64_t* fi = (int64_t*)malloc(sizeof(int64_t) * (T ? T : 1));
    int64_t* fo = (int64_t*)malloc(sizeof(int64_t) * (T ? T : 1));
    double* fw = (double*)malloc(sizeof(double) * (T ? T : 1));
    std::vector<uint64_t> key(T ? T : 1);
    std::vector<double> val(T ? T : 1);
    std::vector<int32_t> cursor(4 * n_in, 0);
    for (int64_t g = 0; g < chunk; g++) {
        int p = pass_of_chunk(B, g);
        const Pass& P = B.pass[p];
        int64_t local = g - P.chunk_base, row = local / P.cpr, c = local % P.cpr;
        int64_t k0, k1;
        chunk_range(B, P, c, &k0, &k1);
        HFlat f{choff[g], fi, fo, fw, vol.data(), wi};
        (void)walk_chunk(B, P, row, k0, k1, start[g], f);
        if (count[g] > cap) {  // k_scatter: overflowed chunks walk again
            HBucket hb{(uint32_t)choff[g], boff.data(), cursor.data(),
                       key.data(), val.data(), p};
            (void)walk_chunk(B, P, row, k0, k1, start[g], hb);
            restaged++;
            continue;
        }
        for (int32_t t = 0; t < count[g]; t++) {
            const Triple e = stage[stage_index(thread_of_chunk(P, row, c), t, cap)];
            int64_t pos = boff[4 * e.cin + p] + cursor[4 * e.cin + p]++;
            key[pos] = ((uint64_t)(uint32_t)e.cout << 32) | (uint64_t)(choff[g] + t);
            val[pos] = e.a;
        }
    }
    // reduce
    std::vector<int64_t> nnz_row(n_in);
    for (int64_t i = 0; i < n_in; i++) {
        bool lr = false;
        int64_t s = boff[4 * i], n = boff[4 * i + 4] - s;
        for (int64_t q = 0; q < n; q++) val[s + q] = weight_of(val[s + q], vol[i], wi, i);
        nnz_row[i] = bucket_reduce(key.data() + s, val.data() + s, n, &lr);
        if (lr) {
            nlong++;
            std::vector<KV> r(n);
            for (int64_t q = 0; q < n; q++) r[q] = KV{key[s + q], val[s + q]};
            nnz_row[i] = bucket_reduce_long(r.data(), n);
            for (int64_t q = 0; q < nnz_row[i]; q++) {
                key[s + q] = r[q].key;
                val[s + q] = r[q].val;
            }
        }
    }
    int64_t nnz = 0;
    for (int64_t i = 0; i < n_in; i++) nnz += nnz_row[i];
    int64_t* ci = (int64_t*)malloc(sizeof(int64_t) * (nnz ? nnz : 1));
    int64_t* co = (int64_t*)malloc(sizeof(int64_t) * (nnz ? nnz : 1));
    double* cw = (double*)malloc(sizeof(double) * (nnz ? nnz : 1));
    int64_t q = 0;
    for (int64_t i = 0; i < n_in; i++)
        for (int64_t a = 0; a < nnz_row[i]; a++) {
            ci[q] = i;
            co[q] = (int64_t)key[boff[4 * i] + a];
            cw[q] = val[boff[4 * i] + a];
            q++;
        }
    *flat_in = fi;
    *flat_out = fo;
    *flat_w = fw;
    *n_flat = T;
    *csr_in = ci;
    *csr_out = co;
    *csr_w = cw;
    stats[0] = chunk;
    stats[1] = fixups;
    stats[2] = anomalies;
    stats[3] = nlong;
    stats[4] = failed;
    stats[5] = restaged;
    stats[6] = map_errors;
    (void)n_out;
    return nnz;
}

void rge_free(void* p) { free(p); }

// sort_network<N> (N = 8 or 16) on each of `count` rows of N keys, with the
// keys' positions as values.
void rge_sort_network(uint64_t* keys, double* vals, int64_t count, int n)
{
    for (int64_t r = 0; r < count; r++) {
        if (n == 8)
            sort_network<8, uint64_t>(keys + 8 * r, vals + 8 * r);
        else
            sort_network<16, uint64_t>(keys + 16 * r, vals + 16 * r);
    }
}

double rge_reduceat_segment(const double* a, int64_t n)
{
    return reduceat_segment(a, n);
}

double rge_reduceat_segment_inplace(double* a, int64_t n, int64_t stride)
{
    return n == 1 ? a[0] : a[0] + pairwise_sum_inplace(a + stride, n - 1, stride);
}

// bucket_reduce_sorted on one sorted bucket (entry j at keys[j * stride]):
// returns the count of unique outputs, or of runs when *long_run is set.
int64_t rge_bucket_reduce(uint64_t* keys, double* vals, int64_t n, int64_t stride,
                          int32_t* long_run)
{
    bool lr = false;
    const int64_t m = bucket_reduce_sorted(keys, vals, n, &lr, stride);
    *long_run = lr ? 1 : 0;
    return m;
}

void rge_cell_area(const double* x, const double* y, int64_t nx, int64_t ny,
                   double* out)
{
    for (int64_t c = 0; c < (nx - 1) * (ny - 1); c++)
        out[c] = cell_area(x, y, ny, c / (ny - 1), c % (ny - 1));
}

int64_t rge_index_secant(const double* x, const double* y, int64_t nx,
                         int64_t ny, double px, double py)
{
    return index_secant(x, y, nx, ny, px, py);
}

// The device-range layout of the graph workspace (rg_layout.hpp) over a
// trace of requests: ops[q] > 0 allocates that many bytes (its offset goes
// to off[q]), ops[q] < 0 frees the allocation made by op -ops[q] - 1.
// Returns the peak.
int64_t rge_layout_trace(const int64_t* ops, int64_t n, int64_t* off)
{
    OffsetLayout L;
    for (int64_t q = 0; q < n; q++) {
        if (ops[q] > 0) {
            off[q] = (int64_t)L.take((size_t)ops[q]);
        } else {
            L.give((size_t)off[-ops[q] - 1]);
            off[q] = -1;
        }
    }
    return (int64_t)L.peak;
}

// The row start's winding number over a boundary polygon, by every edge
// (skip = 0) or skipping the runs of 64 edges whose vertex y range excludes
// py, as k_row_start does (skip = 1).
double rge_winding(const double* bpx, const double* bpy, int64_t nbp, double px, double py,
                   int skip)
{
    double w = 0;
    const int64_t nb = (nbp + 63) / 64;
    for (int64_t b = 0; b < nb; b++) {
        const int64_t k1 = (b + 1) * 64 < nbp ? (b + 1) * 64 : nbp;
        if (skip) {
            double lo = __builtin_huge_val(), hi = -__builtin_huge_val();
            for (int64_t k = b * 64; k < k1; k++) {
                const int64_t km = k == 0 ? nbp - 1 : k - 1;
                lo = fmin(lo, fmin(bpy[km], bpy[k]));
                hi = fmax(hi, fmax(bpy[km], bpy[k]));
            }
            if (py < lo || py > hi) continue;
        }
        for (int64_t k = b * 64; k < k1; k++) {
            const int64_t km = k == 0 ? nbp - 1 : k - 1;
            w += winding_term(px, py, bpx[km], bpy[km], bpx[k], bpy[k]);
        }
    }
    return w;
}

// The cell search on separate x, y arrays and on the interleaved copy, for
// n points: both results per point (they must be equal).
void rge_secant_pair(const double* x, const double* y, int64_t nx, int64_t ny,
                     const double* px, const double* py, int64_t n, int64_t* a, int64_t* b)
{
    std::vector<double2> xy(nx * ny);
    for (int64_t q = 0; q < nx * ny; q++) {
        xy[q].x = x[q];
        xy[q].y = y[q];
    }
    for (int64_t q = 0; q < n; q++) {
        a[q] = index_secant(x, y, nx, ny, px[q], py[q]);
        b[q] = index_secant_xy(xy.data(), nx, ny, px[q], py[q]);
    }
}

}  // extern "C"
