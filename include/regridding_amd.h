/*
 * regridding_amd.h -- C ABI of the MI355X conservative regridding library
 * (librg_hip.so, built for gfx950).
 *
 * Plain pointers and sizes only.  Unless stated otherwise every pointer
 * argument is DEVICE memory on the current device, every function returns an
 * rg_status (0 = success) and, on failure, leaves a message retrievable with
 * rg_last_error() on the calling thread.  `stream` is a hipStream_t (NULL
 * selects the null stream); work is enqueued on it in order.  Cell arrays
 * are flat, row-major over the two resampled axes, exactly like the flat
 * cell indices of the reference (_weights_conservative_2d/_arrays.py:70-91).
 *
 * Each entry point names the reference function it replaces
 * (paths relative to the reference repository's regridding/ package).
 */
#ifndef REGRIDDING_AMD_H
#define REGRIDDING_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum rg_status {
    RG_OK = 0,
    RG_ERR_ARGUMENT = 1,  /* invalid sizes / null pointers (the reference raises
                             ValueError before its compiled call) */
    RG_ERR_HIP = 2,       /* a HIP runtime call failed */
    RG_ERR_MEMORY = 3,    /* device allocation failed */
    RG_ERR_GEOMETRY = 4,  /* the sweep did not terminate (exactly degenerate
                             corner crossings, where the reference loops) */
    RG_ERR_INJECTED = 5,  /* failure forced through rg_inject_failure */
    RG_ERR_CAPACITY = 6   /* a captured graph's replay exceeded its fixed
                             capacities (rg_graph_status) */
};

/* Opaque device-resident weights of one 2-D element. */
typedef struct rg_weights* rg_weights_t;

/* Tuning of the build.  Zero fields select the defaults. */
typedef struct rg_tuning {
    int32_t edges_per_lane; /* sweep edges walked by one lane (default 16) */
    int32_t block;          /* threads per block of the walk kernels (128; a
                               multiple of 64, at most 512) */
    int32_t flags;          /* RG_TUNE_TIMINGS: record per-phase times;
                               RG_TUNE_GUESS_OUTSIDE, RG_TUNE_ROW_WAVES: test
                               hooks */
    int32_t stage_per_edge; /* triples staged per sweep edge by the counting
                               walk (default 8); chunks that emit more are
                               walked a second time */
} rg_tuning;

#define RG_TUNE_TIMINGS 1
/* Test hook: start every chunk after a row's first from "outside" instead of
 * the cell-search guess, so the fix-up re-walks most chunks; the result is
 * the same bits, only slower. */
#define RG_TUNE_GUESS_OUTSIDE 2
/* Test hook: the row starts take 64 rows per wavefront (one secant search
 * per lane) whatever the row count, where a build with few rows takes one
 * row per wavefront; the result is the same bits. */
#define RG_TUNE_ROW_WAVES 4

/* Counters and (optionally) phase times of a build. */
typedef struct rg_build_stats {
    int64_t n_triples;      /* uncoalesced (in, out, w) triples emitted */
    int64_t n_chunks;       /* lanes of the walk */
    int64_t n_fixups;       /* chunks re-walked by the consistency pass */
    int64_t n_subsegments;  /* inside sub-segments (the FP64 work unit) */
    int64_t n_search_anomalies; /* row starts inside the boundary polygon
                                   that the cell search did not place */
    int64_t n_restaged;     /* chunks whose triples overflowed the staging */
    int64_t n_long_cells;   /* input cells reduced in global memory (more
                               records than the LDS share, or a run of more
                               than 129 terms) */
    /* with RG_TUNE_TIMINGS, stream times (HIP events) of: prepare (areas,
     * static-grid tables), start (row and chunk start states), count (the
     * counting walk kernel), fixup (fix-up walk, scans and the host wait
     * for the totals), emit (the scatter kernel), reduce (the bucket
     * reduction kernels), transpose (both CSR layouts), total */
    double ms_prepare, ms_start, ms_count, ms_fixup, ms_emit, ms_reduce,
        ms_transpose, ms_total;
} rg_build_stats;

/* Library / device. */
const char* rg_last_error(void);
int rg_version(void);
int rg_init(int device);
int rg_device_count(int* count);
int rg_mem_info(size_t* free_bytes, size_t* total_bytes);
void rg_trim(void); /* release cached device blocks */

/* Device memory plumbing for hosts without their own allocator (ctypes). */
int rg_malloc(void** ptr, size_t bytes);
int rg_free(void* ptr);
int rg_memcpy_htod(void* dst, const void* src, size_t bytes, void* stream);
int rg_memcpy_dtoh(void* dst, const void* src, size_t bytes, void* stream);
int rg_memset(void* dst, int value, size_t bytes, void* stream);
/* Pinned (page-locked) host memory, for asynchronous transfers. */
int rg_host_alloc(void** ptr, size_t bytes);
int rg_host_free(void* ptr);
int rg_stream_create(void** stream);
int rg_stream_destroy(void* stream);
int rg_stream_sync(void* stream);
int rg_device_sync(void);

/* HIP event timers (for hosts that time kernels on their own stream). */
int rg_event_create(void** event);
int rg_event_destroy(void* event);
int rg_event_record(void* event, void* stream);
int rg_event_elapsed_ms(void* start, void* stop, float* ms); /* waits on stop */

/* Plain device copy of n doubles through a streaming kernel (no reference
 * counterpart): the known-size reference for hardware traffic counters. */
int rg_copy_f64(const double* src, double* dst, int64_t n, void* stream);

/* Replaces _weights_conservative_2d/_grids.py:50 grid_volume: signed
 * shoelace area of every cell of an (nx, ny) vertex grid, accumulated in the
 * reference's order.  area has (nx-1)*(ny-1) entries. */
int rg_cell_area_2d(const double* x, const double* y, int64_t nx, int64_t ny,
                    double* area, void* stream);

/* Replaces _weights_conservative_2d/_weights_conservative_2d.py:80
 * weights_conservative_2d followed by _weights_arrays.py:44 _coalesce for one
 * element: builds the coalesced sparse weights between an input grid
 * (nxi, nyi) and an (already perturbed) output grid (nxo, nyo).  wi is the
 * optional (nxi-1)*(nyi-1) input weighting or NULL.  tuning may be NULL.
 * The returned handle owns its device memory; release with
 * rg_weights_free. */
int rg_build_conservative_2d(const double* xi, const double* yi, int64_t nxi,
                             int64_t nyi, const double* xo, const double* yo,
                             int64_t nxo, int64_t nyo, const double* wi,
                             const rg_tuning* tuning, rg_weights_t* out,
                             void* stream);

/* One grid pair of a batched build: the arguments of
 * rg_build_conservative_2d (device pointers). */
typedef struct rg_pair_desc {
    const double* xi;
    const double* yi;
    int64_t nxi, nyi;
    const double* xo;
    const double* yo;
    int64_t nxo, nyo;
    const double* wi; /* (nxi-1)*(nyi-1) input weighting, or NULL */
} rg_pair_desc;

/* Replaces the per-element loop of
 * _weights/_weights_conservative.py:122-139, which calls
 * weights_conservative_2d (and then _coalesce) once per orthogonal element:
 * builds npairs independent grid pairs, of any mix of sizes, together --
 * every phase is one launch over the concatenated work of all pairs.
 * out[k] receives the handle of pairs[k], identical (bit for bit) to what
 * rg_build_conservative_2d returns for that pair alone.  The handles share
 * device buffers, released with the last of them (rg_weights_free each).
 * In rg_weights_stats of a batched handle, n_triples and n_chunks are the
 * pair's; the other counters and the phase times are the batch's. */
int rg_build_conservative_2d_batched(const rg_pair_desc* pairs, int64_t npairs,
                                     const rg_tuning* tuning, rg_weights_t* out,
                                     void* stream);

/* Autotuner of the (batched) build; no reference counterpart.  Builds
 * `pairs` once per candidate (edges_per_lane[a], block[b]) -- reps timed
 * runs each after one untimed run, HIP events on `stream` -- and returns
 * the fastest setting in *best (stage_per_edge and flags from `base`, which
 * may be NULL) and its median time in *best_ms.  times, if not NULL,
 * receives the n_epl x n_block median times (ms), row-major.  Results of
 * a build do not depend on the tuning. */
int rg_tune_build(const rg_pair_desc* pairs, int64_t npairs,
                  const int32_t* edges_per_lane, int n_epl,
                  const int32_t* block, int n_block, int reps,
                  const rg_tuning* base, rg_tuning* best, double* best_ms,
                  double* times, void* stream);

/* Sizes and build counters of a handle. */
int rg_weights_info(rg_weights_t w, int64_t* n_in, int64_t* n_out,
                    int64_t* nnz);
int rg_weights_stats(rg_weights_t w, rg_build_stats* stats);

/* Replaces the output arrays of _weights_arrays.py:68-73: writes the nnz
 * coalesced triples sorted by (input, output) into device arrays. */
int rg_weights_export(rg_weights_t w, int64_t* indices_input,
                      int64_t* indices_output, double* values, void* stream);

/* Replaces the TypedList marshalling of _regrid_from_weights.py:139-141:
 * adopts nnz (input, output, weight) triples in ANY order (device arrays),
 * e.g. unpickled or transposed weights.  Applying the handle reproduces the
 * reference's sequential scatter in storage order bit for bit. */
int rg_weights_import(const int64_t* indices_input,
                      const int64_t* indices_output, const double* values,
                      int64_t nnz, int64_t n_in, int64_t n_out,
                      rg_weights_t* out, void* stream);

int rg_weights_free(rg_weights_t w);

/* Declares the 2-D cell shapes of imported weights (built weights know
 * them): the apply walks its outputs in 2-D tiles for cache locality.  Has
 * no effect on results. */
int rg_weights_set_shape(rg_weights_t w, int64_t ncx_in, int64_t ncy_in,
                         int64_t ncx_out, int64_t ncy_out);

/* Launch shape of the apply kernels (process-wide; 0 keeps the current
 * value): tile_w output cells along the last axis per tile row of a
 * 256-cell tile for the FP64 forward apply, twice that (up to 256) for FP32
 * values and for the adjoint (default 16; negative: same widths, staged
 * input patches off),
 * images_per_pass images accumulated per pass over the weights (1, 2, 4, 8
 * or 16), lds_kib the staged patch size per block (KiB, 1..64) above which
 * fewer images are staged per pass, and rounds (1: one block per tile loops
 * over the image groups with its CSR entries in registers; 0: one block per
 * tile and image group; -1 keeps the current choice).  Has no effect on
 * results. */
int rg_set_apply_tuning(int tile_w, int images_per_pass, int lds_kib, int rounds);

/* Replaces _regrid/_regrid_from_weights.py:166 _regrid_from_weights for
 * `batch` elements that share one weight matrix: values_input is
 * (batch, n_in), values_output (batch, n_out), both contiguous.
 * transpose=0 applies the weights (out[o] = sum_i w[i,o] in[i]);
 * transpose=1 applies the transposed weights, which is what the reference
 * computes for regrid_from_weights(*transpose_weights(weights), ...)
 * (_weights_transposed.py:13-50).  Each output is accumulated sequentially
 * in the reference's storage order, so the FP64 result is bit-identical. */
int rg_apply_f64(rg_weights_t w, int transpose, const double* values_input,
                 double* values_output, int64_t batch, void* stream);

/* Single-precision variant: FP32 values, FP32 weights (rounded from the
 * FP64 weights), FP32 accumulation in the same order.  No reference
 * counterpart (the reference always accumulates in float64). */
int rg_apply_f32(rg_weights_t w, int transpose, const float* values_input,
                 float* values_output, int64_t batch, void* stream);

/* One apply of a batched apply call: as the arguments of rg_apply_f64 /
 * rg_apply_f32 (values are float64 for the f64 call, float32 for f32). */
typedef struct rg_apply_job {
    rg_weights_t weights;
    int32_t transpose;
    const void* values_input;
    void* values_output;
    int64_t batch;
} rg_apply_job;

/* Replaces the loop of _regrid/_regrid_from_weights.py:172-182 over
 * orthogonal elements when the elements carry different weights (the
 * object array returned by weights() for curvilinear grids that vary along
 * the orthogonal axes): every job's apply in one launch.  Each job's result
 * is bit-identical to its own rg_apply_f64 / rg_apply_f32 call. */
int rg_apply_batched_f64(const rg_apply_job* jobs, int64_t njobs, void* stream);
int rg_apply_batched_f32(const rg_apply_job* jobs, int64_t njobs, void* stream);

/* Replaces _weights_conservative_2d/_grids.py:356 index_of_point_secant
 * (with its index_of_point_brute fallback, :223), vectorised over n query
 * points: cell_ij receives 2*n int64 values (i, j) per point, or
 * (INT64_MAX, INT64_MAX) for points outside the grid. */
int rg_find_cells_2d(const double* x, const double* y, int64_t nx, int64_t ny,
                     const double* px, const double* py, int64_t n,
                     int64_t* cell_ij, void* stream);

/* No counterpart in the reference: rg_find_cells_2d for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and results, bit for
 * bit; it calls no allocator.  Its 8 bytes of scratch per point live in a
 * fixed device array of RG_FIND_CELLS_NOALLOC_MAX points per GPU (16 MiB,
 * held until the process ends); larger n fails with RG_ERR_ARGUMENT (search
 * in slices).  Every call on one GPU shares that array: calls that overlap
 * on one GPU race, so issue them on one stream. */
#define RG_FIND_CELLS_NOALLOC_MAX (1 << 21)
int rg_find_cells_2d_noalloc(const double* x, const double* y, int64_t nx,
                             int64_t ny, const double* px, const double* py,
                             int64_t n, int64_t* cell_ij, void* stream);

#define RG_FIND_BRUTE 0
#define RG_FIND_SEARCHSORTED 1

/* Replaces _find_indices/_find_indices_brute.py:25 and
 * _find_indices_searchsorted.py:25 (_find_indices_*_1d): for each of `batch`
 * elements, the cell of the m-vertex input grid containing each of the n
 * output points, or fill_value.  x_input is (batch, m), x_output (batch, n),
 * result (batch, n). */
int rg_find_indices_1d(const double* x_input, int64_t m,
                       const double* x_output, int64_t n, int64_t batch,
                       int method, int64_t fill_value, int64_t* result,
                       void* stream);

/* Replaces weights_conservative_1d (_weights/_weights_conservative_1d/
 * _weights_conservative_1d.py:12-205) followed by _coalesce
 * (_weights_arrays.py:44) for D elements, in two calls.  x_input is
 * (D, n_input) vertices, x_output (D, n_output).  The count call fills
 * counts (D*(n_input-1) int32: triples per input cell) and offsets
 * (D*(n_input-1)+1 int64, their exclusive scan) and returns the total; the
 * emit call writes the triples (the reference's (possibly negative) indices)
 * in coalesced order: element by element, sorted by (input, output).
 * weights_input is (D, n_input-1) or NULL. */
int rg_conservative_1d_count(const double* x_input, int64_t n_input,
                             const double* x_output, int64_t n_output, int64_t D,
                             int32_t* counts, int64_t* offsets, int64_t* total,
                             void* stream);
int rg_conservative_1d_emit(const double* x_input, int64_t n_input,
                            const double* x_output, int64_t n_output, int64_t D,
                            const double* weights_input, const int32_t* counts,
                            const int64_t* offsets, int64_t* indices_input,
                            int64_t* indices_output, double* values, void* stream);

/* Replaces _weights_transposed.py:309-325 _cell_volume_1d (cell_length of
 * _weights_conservative_1d/_grids.py:11-37): the signed length of every cell
 * of D elements of n vertices; length is (D, n-1). */
int rg_cell_length_1d(const double* x, int64_t n, int64_t D, double* length,
                      void* stream);

/* Replaces _weights_multilinear.py:148-206 (_weights_from_indices_multilinear_1d
 * with the clamping of :95-118) for D elements: from the searchsorted indices
 * of the output vertices (D, n_output), writes 2*n_output triples per element
 * in the reference's storage order; bounds_nan = 1 marks outside points NaN.
 * n_outside receives the number of outside points (for bounds="raise").
 * weights_input is (D, n_input) or NULL. */
int rg_multilinear_1d(const double* x_input, int64_t n_input, const double* x_output,
                      int64_t n_output, int64_t D, const int64_t* indices,
                      const double* weights_input, int32_t bounds_nan,
                      int64_t* indices_input, int64_t* indices_output, double* values,
                      int64_t* n_outside, void* stream);

/* Test hook (no reference counterpart): make the next call that reaches the
 * named site ("build", "apply", "alloc") fail with RG_ERR_INJECTED.  NULL
 * clears pending injections. */
int rg_inject_failure(const char* site);

/* ---- one hipGraph for a batched build and apply ---------------------
 *
 * No counterpart in the reference: the batched build (above) followed by
 * the FP64 apply of one job per pair, captured once into a hipGraph and
 * replayed with one launch.  A replay reads the coordinates and values at
 * the same device addresses as at creation (their contents may change;
 * the shapes may not) and writes the outputs there.  rg_graph_create builds
 * once to measure the data-dependent sizes, sizes a workspace with `slack`
 * (0.25: a quarter more triples and weights than measured), then captures.
 * A replay whose data exceed a capacity (or whose sweep fails) leaves
 * incomplete results and sets rg_graph_status's flags: rebuild with more
 * slack.  The graph's buffers are its own: build handles with
 * rg_graph_weights (a copy of pair k's weights from the last replay;
 * RG_ERR_CAPACITY when that replay set flags).  Status and weights need a
 * launch first.  Replays on one graph share its buffers: they must not
 * overlap (launch them on one stream). */
typedef struct rg_graph* rg_graph_t;

typedef struct rg_graph_job {
    const double* values_input; /* (batch, n_in) of the pair, device */
    double* values_output;      /* (batch, n_out), device */
    int64_t batch;              /* 0: no apply for this pair */
} rg_graph_job;

/* flags of rg_graph_status */
#define RG_GRAPH_TRIPLES_CAPACITY 1
#define RG_GRAPH_OVERFLOW_CAPACITY 2
#define RG_GRAPH_WEIGHTS_CAPACITY 4
#define RG_GRAPH_SWEEP_FAILED 8
#define RG_GRAPH_MISMATCH 16

int rg_graph_create(const rg_pair_desc* pairs, int64_t npairs, const rg_tuning* tuning,
                    const rg_graph_job* jobs, double slack, rg_graph_t* out, void* stream);
int rg_graph_launch(rg_graph_t g, void* stream);
/* waits for `stream` and reads the flags of the last replay (0: complete) */
int rg_graph_status(rg_graph_t g, int32_t* flags, void* stream);
int rg_graph_weights(rg_graph_t g, int64_t k, rg_weights_t* out, void* stream);
int rg_graph_capacity(rg_graph_t g, int64_t* triples, int64_t* overflow, int64_t* weights,
                      int64_t* workspace_bytes);
int rg_graph_destroy(rg_graph_t g);

/* Replaces _fill/_gauss_seidel.py:83-139 (_fill_gauss_seidel_2d): num_iterations
 * red-black Gauss-Seidel sweeps, in place, over T elements of (ny, nx) values
 * with periodic neighbours, updating the cells where where != 0; the same
 * update order as the reference's sequential sweep. */
int rg_fill_gauss_seidel_2d(double* a, const uint8_t* where, int64_t T, int64_t ny,
                            int64_t nx, int64_t num_iterations, void* stream);

/* Replace _interp_ndarray.py:172-187 and :190-205
 * (_ndarray_linear_interpolation_1d / _2d): D elements of a (n) or (nx, ny)
 * evaluated at m fractional indices each (x, or x and y), clamped to the
 * last cell; out is (D, m). */
int rg_interp_linear_1d(const double* a, int64_t n, const double* x, int64_t m,
                        int64_t D, double* out, void* stream);
int rg_interp_linear_2d(const double* a, int64_t nx, int64_t ny, const double* x,
                        const double* y, int64_t m, int64_t D, double* out, void* stream);

/* ---- serving loop (no reference counterpart) ------------------------------
 *
 * A request is one regrid job: a grid pair (or the cache key of weights
 * built earlier) and a stack of value arrays.  Submitted requests queue in
 * arrival order; worker threads, each with its own stream, repeatedly take
 * the longest prefix of the queue that fits the free device memory (each
 * request's footprint is estimated from its sizes), build the weights of
 * all cache misses of that group in one batched build, apply every request
 * of the group in one batched apply, and copy the results back.  Groups of
 * different workers overlap on the device; a request that arrives while
 * others run joins the next group a worker forms.  Built weights stay on the
 * device, keyed by the caller's grid hash (LRU, bounded by cache_bytes). */

#define RG_F64 0
#define RG_F32 1

typedef struct rg_serve_config {
    int32_t workers;       /* groups in flight, one stream each (default 3) */
    int32_t max_group;     /* most requests in one group (default 64) */
    int64_t reserve_bytes; /* device memory admission leaves free (4 GiB) */
    int64_t cache_bytes;   /* weight cache capacity (default 32 GiB) */
    rg_tuning tuning;      /* tuning of the batched builds */
} rg_serve_config;

typedef struct rg_serve_request {
    uint64_t key;      /* weight-cache key (equal keys: equal weights); 0: not cached */
    const double* xi;  /* host vertex grids, already perturbed (read on a miss) */
    const double* yi;
    int64_t nxi, nyi;
    const double* xo;
    const double* yo;
    int64_t nxo, nyo;
    const double* wi;       /* host input weighting or NULL */
    int32_t dtype;          /* RG_F64 or RG_F32 */
    int32_t transpose;      /* 1: apply the adjoint (output cells -> input cells) */
    const void* values;     /* host (batch, n_in), or (batch, n_out) if transpose */
    void* result;           /* host (batch, n_out), or (batch, n_in) if transpose */
    int64_t batch;
    int64_t arrival_ns;     /* rg_clock_ns() time of arrival (0: at submission) */
} rg_serve_request;

typedef struct rg_serve_stats {
    int64_t submitted, completed, failed;
    int64_t cache_hits, cache_misses, builds, groups;
    int64_t max_group_seen;
    int64_t peak_admitted_bytes;
    int64_t cache_entries, cache_bytes;
} rg_serve_stats;

typedef struct rg_server* rg_server_t;

/* Monotonic clock (ns) of arrival and completion times. */
int64_t rg_clock_ns(void);
int rg_serve_create(const rg_serve_config* config, rg_server_t* out);
/* Thread-safe.  The request's host buffers must stay valid until
 * rg_serve_wait returns for its ticket. */
int rg_serve_submit(rg_server_t server, const rg_serve_request* request,
                    uint64_t* ticket);
/* Blocks until the request is done; returns its status (rg_last_error holds
 * its message on failure) and its completion time.  Each ticket is waited
 * for once. */
int rg_serve_wait(rg_server_t server, uint64_t ticket, int64_t* done_ns);
int rg_serve_stats_get(rg_server_t server, rg_serve_stats* stats);
/* Finishes every submitted request, then stops the workers and frees the
 * cache. */
int rg_serve_destroy(rg_server_t server);

#ifdef __cplusplus
}
#endif

#endif /* REGRIDDING_AMD_H */
