// rg_internal.hpp -- host-side plumbing shared by the HIP translation units:
// status codes and the thread-local error string, a caching device
// allocator, device-wide exclusive scans, and the weights object behind the
// opaque rg_weights_t handle.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <memory>
#include <string>
#include <vector>

#include "../../include/regridding_amd.h"

namespace rg {

// ---- errors --------------------------------------------------------------

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define RG_HIP(expr)                                                        \
    do {                                                                    \
        hipError_t e_ = (expr);                                             \
        if (e_ != hipSuccess)                                               \
            return ::rg::fail(RG_ERR_HIP, std::string(#expr) + ": " +      \
                                              hipGetErrorString(e_));       \
    } while (0)

#define RG_TRY(expr)                \
    do {                            \
        int s_ = (expr);            \
        if (s_ != RG_OK) return s_; \
    } while (0)

// ---- device memory -------------------------------------------------------

// Caching allocator: freed blocks are kept per (device, stream) and reused
// only by later requests on the same stream, so a block released while
// kernels that read it are still queued is never handed to concurrent work;
// repeated builds and applies do not pay hipMalloc/hipFree.  rg_trim()
// synchronises the device and releases the cache.
int dev_alloc(void** p, size_t bytes, hipStream_t s);
void dev_free(void* p);
void dev_trim();
// Arena mode of the allocator for the calling host thread (see rg_runtime.hip).
void arena_begin(void* base, size_t cap);
size_t arena_end();
void arena_measure_begin();
size_t arena_measure_end();

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { reset(); }
    int alloc(size_t count, hipStream_t s)
    {
        reset();
        n = count;
        if (count == 0) return RG_OK;
        return dev_alloc((void**)&p, count * sizeof(T), s);
    }
    void reset()
    {
        if (p) dev_free(p);
        p = nullptr;
        n = 0;
    }
    T* release()
    {
        T* q = p;
        p = nullptr;
        n = 0;
        return q;
    }
};

// Pinned host scratch for small synchronous device->host reads.
int read_scalars(const void* dev, void* host, size_t bytes, hipStream_t s);
// The same read in two steps: _begin queues the copy on s, _end waits for
// that copy alone (work queued on s in between keeps running) and returns
// the bytes.  One such read per host thread at a time.
int read_scalars_begin(const void* dev, size_t bytes, hipStream_t s);
int read_scalars_end(void* host, size_t bytes);

// Device address, on the current device, of the fixed scratch array of
// rg_find_cells_2d_noalloc (looked up once per device and kept).
int find_cells_scratch(int64_t** out);

// ---- scans ---------------------------------------------------------------

// out[0..n] = exclusive prefix sum of in[0..n), out[n] = total.  `in` and
// `out` may not alias.
int scan_i32_to_i64(const int32_t* in, int64_t* out, int64_t n, hipStream_t s);
int scan_i64(const int64_t* in, int64_t* out, int64_t n, hipStream_t s);

// ---- fault injection (host-forced error paths for tests) -----------------

bool injected_failure(const char* site);

// ---- launch geometry -----------------------------------------------------

inline unsigned grid_for(int64_t n, int block)
{
    int64_t g = (n + block - 1) / block;
    return (unsigned)(g < 1 ? 1 : g);
}

}  // namespace rg

// The object behind rg_weights_t: sparse weights of one 2-D element, stored
// twice on the device.  CSR by input cell (columns = output cells ascending)
// is the reference's storage order (sorted by (input, output)); CSR by
// output cell (columns = input cells ascending) drives the forward apply.
// Device buffers shared by the handles of one batched build: the handles'
// CSR arrays are slices of them, freed with the last handle.
namespace rg {
// Largest k in [0, n) with base[k] <= x (base ascending, base[0] <= x).
__device__ inline int seg_of(const int64_t* __restrict__ base, int n, int64_t x)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (base[mid] <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}
}  // namespace rg

namespace rg {
// One job of a batched apply (rg_apply.hip): CSR rows of one weight set
// applied to `batch` value arrays; cut into blocks of (tile of rows) x
// images.  The rows are the cells of a (gx, gy) grid and a tile is tw x th
// of them (job_tiles sets the tiling).
struct JobDesc {
    const int32_t* rowptr;
    const int32_t* col;
    const void* w;
    const void* x;
    void* y;
    int64_t nrows, ncols, batch;
    int64_t block_base;  // first block of this job
    int64_t nrt;         // row tiles of the job
    int64_t gy, ntx;     // grid width (rows per grid row); tiles per tile row
    int32_t tw, th;      // tile width and height (tw * th = block size)
};
constexpr int JOB_DB = 8;  // images per block
// Tiles the job's rows, the cells of a (gx, gy) grid (gx * gy == nrows; any
// other shape is taken as one long row), into blocks; sets gy, ntx, tw, th
// and nrt.
void job_tiles(JobDesc& d, int64_t gx, int64_t gy, int64_t value_bytes);
// Launches the FP64 batched apply over a device job table (block prefix
// `prefix`, `nblocks` blocks in all); every block returns at once when
// *abort_flag is nonzero (abort_flag may be null).
int apply_jobs_device(const JobDesc* jobs, const int64_t* prefix, int njobs, int64_t nblocks,
                      const int32_t* abort_flag, hipStream_t s);
}  // namespace rg

struct rg_arena {
    std::vector<void*> bufs;
    ~rg_arena();
};

struct rg_weights {
    int device = 0;
    int64_t n_in = 0, n_out = 0, nnz = 0;
    // 2-D cell shapes of the two grids (rows of the CSRs are flat cells);
    // (1, n) when unknown (imported weights without rg_weights_set_shape)
    int64_t shape_in[2] = {1, 0}, shape_out[2] = {1, 0};
    int32_t* rowptr_in = nullptr;   // n_in + 1
    int32_t* col_in = nullptr;      // nnz, output cell of each entry
    double* val_in = nullptr;       // nnz
    int32_t* rowptr_out = nullptr;  // n_out + 1
    int32_t* col_out = nullptr;     // nnz, input cell of each entry
    double* val_out = nullptr;      // nnz
    float* val_in_f32 = nullptr;    // lazily made single-precision copies
    float* val_out_f32 = nullptr;
    // apply plans (see rg_apply.hip), one per direction, made on first use
    // One per direction and value size (index 2 * dir + (4-byte values)):
    // FP32 applies use tiles twice as wide as FP64 ones.
    struct Plan {
        int64_t tile_w = 0;        // tile shape the plan was made for
        int32_t* box = nullptr;    // per tile: i0, j0, rows, cols of the patch
        uint16_t* off = nullptr;   // per CSR entry: offset inside its patch
        int64_t ntiles = 0;
        int64_t max_cells = 0;     // largest patch (cells)
    } plan[4];
    rg_build_stats stats = {};
    // set when the six CSR arrays are slices of a batched build's buffers
    std::shared_ptr<rg_arena> arena;
    // The stream the handle was made on, and every other stream that has
    // applied it with an event recorded after its latest apply there: the
    // device memory goes back to the pool of the stream it was allocated
    // on, so release first waits for applies still queued elsewhere.
    hipStream_t home = nullptr;
    std::vector<std::pair<hipStream_t, hipEvent_t>> uses;
};

namespace rg {
// Notes an apply of W on stream s (after its launch): records s's event
// unless s is the handle's own stream.
int note_use(rg_weights* W, hipStream_t s);
}  // namespace rg
