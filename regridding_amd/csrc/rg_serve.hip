// rg_serve.hip -- the serving loop: continuous batching of regrid requests.
//
// Requests queue in arrival order.  Each worker thread owns a stream and
// loops: take the longest prefix of the queue whose estimated device
// footprint fits the admission budget (requests whose weights another
// worker is building right now stay queued until that build lands), build
// the cache misses of the group in one batched build, apply every request
// of the group in one batched apply per precision, copy the results back,
// wake the waiters.  With several workers, one group's transfers and
// kernels overlap another's on the device.
//
// The weight cache maps the caller's key to a handle; entries in use by a
// group are pinned, and unpinned entries are evicted least recently used
// first once the cache exceeds its byte budget.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "rg_internal.hpp"

namespace rg {
int build_batched(const rg_pair_desc* pairs, int64_t npairs,
                  const rg_tuning* tuning, rg_weights** out, hipStream_t s);
void weights_release(rg_weights* W);
}  // namespace rg

using namespace rg;

namespace {

int64_t now_ns()
{
    return std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

struct Req {
    rg_serve_request r;
    uint64_t ticket = 0;
    int64_t footprint = 0;
    bool done = false;
    bool failed_alone = false;  // rejected by itself; the group went on
    int status = RG_OK;
    std::string error;
    int64_t done_ns = 0;
};

struct CacheEntry {
    rg_weights* w = nullptr;  // null while being built
    int64_t bytes = 0;
    int pins = 0;
    std::list<uint64_t>::iterator lru;
};

int64_t cells_in(const rg_serve_request& r) { return (r.nxi - 1) * (r.nyi - 1); }
int64_t cells_out(const rg_serve_request& r) { return (r.nxo - 1) * (r.nyo - 1); }

// Device bytes a request holds while in a group: its values and results,
// and for a build its coordinates plus the build's working set (about 400
// bytes per sweep edge: staging, buckets, records, both CSR layouts).
int64_t footprint(const rg_serve_request& r, bool build)
{
    const int64_t es = r.dtype == RG_F32 ? 4 : 8;
    int64_t b = r.batch * (cells_in(r) + cells_out(r)) * es;
    if (build) {
        const int64_t vi = r.nxi * r.nyi, vo = r.nxo * r.nyo;
        b += 16 * (vi + vo) + 400 * 2 * (vi + vo);
    }
    return b;
}

int64_t weights_bytes(const rg_weights* w)
{
    return w->nnz * 24 + (w->n_in + w->n_out + 2) * 4;
}

}  // namespace

struct rg_server {
    rg_serve_config cfg;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Req*> queue;
    std::unordered_map<uint64_t, std::unique_ptr<Req>> reqs;  // by ticket
    std::unordered_map<uint64_t, CacheEntry> cache;
    std::list<uint64_t> lru;  // front = most recent
    int64_t cache_bytes = 0;
    int64_t budget = 0;       // admission budget (bytes)
    int64_t admitted = 0;     // footprint of the groups in flight
    uint64_t next_ticket = 1;
    bool stop = false;
    int device = 0;
    rg_serve_stats st = {};
    std::vector<std::thread> workers;

    void evict_locked()
    {
        auto it = lru.end();
        while (cache_bytes > cfg.cache_bytes && it != lru.begin()) {
            --it;
            auto ce = cache.find(*it);
            if (ce == cache.end() || ce->second.pins > 0 || !ce->second.w) continue;
            cache_bytes -= ce->second.bytes;
            weights_release(ce->second.w);
            cache.erase(ce);
            it = lru.erase(it);
        }
    }

    // Take a group off the queue (caller holds mu).  Returns false when the
    // queue holds nothing runnable.
    bool take_group_locked(std::vector<Req*>& group, std::vector<uint64_t>& build_keys,
                           std::vector<Req*>& build_reqs)
    {
        group.clear();
        build_keys.clear();
        build_reqs.clear();
        int64_t fp = 0;
        for (auto it = queue.begin(); it != queue.end() && (int)group.size() < cfg.max_group;) {
            Req* q = *it;
            const uint64_t key = q->r.key;
            bool miss = true, wait = false, first_build = false;
            if (key) {
                auto ce = cache.find(key);
                if (ce != cache.end()) {
                    if (ce->second.w) {
                        miss = false;
                    } else if (std::find(build_keys.begin(), build_keys.end(), key) ==
                               build_keys.end()) {
                        wait = true;  // another worker is building it
                    } else {
                        miss = false;  // built by this group
                    }
                } else {
                    first_build = true;
                }
            }
            if (wait) {
                ++it;
                continue;
            }
            const bool builds = miss && (key == 0 || first_build);
            const int64_t f = footprint(q->r, builds);
            if (admitted + fp + f > budget && !(group.empty() && admitted == 0)) break;
            fp += f;
            q->footprint = f;
            if (builds) {
                build_reqs.push_back(q);
                if (key) {
                    build_keys.push_back(key);
                    CacheEntry& e = cache[key];  // placeholder: being built
                    lru.push_front(key);
                    e.lru = lru.begin();
                }
            }
            if (key && !builds) {
                CacheEntry& e = cache[key];
                e.pins++;
                if (e.w) {
                    lru.erase(e.lru);
                    lru.push_front(key);
                    e.lru = lru.begin();
                    st.cache_hits++;
                }
            }
            if (builds) st.cache_misses++;
            group.push_back(q);
            it = queue.erase(it);
        }
        admitted += fp;
        st.peak_admitted_bytes = std::max(st.peak_admitted_bytes, admitted);
        return !group.empty();
    }

    void run_group(hipStream_t s, std::vector<Req*>& group,
                   std::vector<uint64_t>& build_keys, std::vector<Req*>& build_reqs);
    void worker();
};

void rg_server::worker()
{
    (void)hipSetDevice(device);
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) s = nullptr;
    std::vector<Req*> group, build_reqs;
    std::vector<uint64_t> build_keys;
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(mu);
            // (a group is taken first, so a stopping server still drains)
            while (!take_group_locked(group, build_keys, build_reqs)) {
                if (stop && queue.empty()) break;
                if (stop) cv_work.wait_for(lk, std::chrono::milliseconds(1));
                else cv_work.wait(lk);
            }
            if (group.empty()) break;
        }
        run_group(s, group, build_keys, build_reqs);
        cv_work.notify_all();  // memory and cache entries may have freed up
    }
    if (s) (void)hipStreamDestroy(s);
}

void rg_server::run_group(hipStream_t s, std::vector<Req*>& group,
                          std::vector<uint64_t>& build_keys,
                          std::vector<Req*>& build_reqs)
{
    int status = RG_OK;
    std::string error;
    auto failed = [&](int code) {
        if (status == RG_OK) {
            status = code;
            error = rg_last_error();
        }
    };
    // ---- builds of the misses: coordinates to the device, one batched build
    std::vector<rg_weights*> built(build_reqs.size(), nullptr);
    std::vector<void*> dev_bufs;
    if (!build_reqs.empty()) {
        std::vector<rg_pair_desc> descs(build_reqs.size());
        for (size_t k = 0; k < build_reqs.size() && status == RG_OK; k++) {
            const rg_serve_request& r = build_reqs[k]->r;
            if (!r.xi || !r.yi || !r.xo || !r.yo) {
                status = fail(RG_ERR_ARGUMENT, "a request without cached weights needs its grids");
                error = rg_last_error();
                break;
            }
            const int64_t vi = r.nxi * r.nyi, vo = r.nxo * r.nyo;
            const int64_t nw = r.wi ? cells_in(r) : 0;
            void* d = nullptr;
            int e = dev_alloc(&d, sizeof(double) * (2 * vi + 2 * vo + nw), s);
            if (e != RG_OK) {
                failed(e);
                break;
            }
            dev_bufs.push_back(d);
            double* p = (double*)d;
            const double* src[5] = {r.xi, r.yi, r.xo, r.yo, r.wi};
            const int64_t cnt[5] = {vi, vi, vo, vo, nw};
            double* dst[5];
            for (int a = 0; a < 5; a++) {
                dst[a] = p;
                if (cnt[a] && hipMemcpyAsync(p, src[a], sizeof(double) * cnt[a],
                                             hipMemcpyHostToDevice, s) != hipSuccess)
                    failed(fail(RG_ERR_HIP, "coordinate upload failed"));
                p += cnt[a];
            }
            rg_pair_desc& dd = descs[k];
            dd.xi = dst[0];
            dd.yi = dst[1];
            dd.nxi = r.nxi;
            dd.nyi = r.nyi;
            dd.xo = dst[2];
            dd.yo = dst[3];
            dd.nxo = r.nxo;
            dd.nyo = r.nyo;
            dd.wi = r.wi ? dst[4] : nullptr;
        }
        if (status == RG_OK) {
            int e = build_batched(descs.data(), (int64_t)descs.size(), &cfg.tuning,
                                  built.data(), s);
            if (e != RG_OK) failed(e);
        }
        // the build synchronised its stream: the coordinates are consumed
        for (void* d : dev_bufs) dev_free(d);
        dev_bufs.clear();
    }
    // publish the built weights
    std::vector<rg_weights*> uncached;  // key-0 builds, freed after the group
    {
        std::lock_guard<std::mutex> lk(mu);
        st.groups++;
        st.max_group_seen = std::max<int64_t>(st.max_group_seen, (int64_t)group.size());
        st.builds += status == RG_OK ? (int64_t)build_reqs.size() : 0;
        for (size_t k = 0; k < build_reqs.size(); k++) {
            const uint64_t key = build_reqs[k]->r.key;
            if (!key) continue;
            CacheEntry& e = cache[key];
            if (status == RG_OK) {
                e.w = built[k];
                e.bytes = weights_bytes(built[k]);
                cache_bytes += e.bytes;
                e.pins++;
            } else {  // drop the placeholder so a later request retries
                lru.erase(e.lru);
                cache.erase(key);
            }
        }
    }
    // ---- the applies: one batched call per precision
    std::vector<rg_apply_job> jobs[2];
    std::vector<std::pair<Req*, void*>> outs;  // request, device result
    if (status == RG_OK) {
        for (Req* q : group) {
            const rg_serve_request& r = q->r;
            rg_weights* w = nullptr;
            if (r.key) {
                std::lock_guard<std::mutex> lk(mu);
                w = cache[r.key].w;
            } else {
                for (size_t k = 0; k < build_reqs.size(); k++)
                    if (build_reqs[k] == q) w = built[k];
                uncached.push_back(w);
            }
            // the caller's buffers are sized from the request's shapes: a
            // cached key whose weights have other sizes (wrong shapes, or a
            // key collision) fails this request alone
            if (w->n_in != cells_in(r) || w->n_out != cells_out(r)) {
                q->failed_alone = true;
                q->status = fail(RG_ERR_ARGUMENT,
                                 "the cached weights of this key map " +
                                     std::to_string(w->n_in) + " -> " +
                                     std::to_string(w->n_out) +
                                     " cells, the request's shapes " +
                                     std::to_string(cells_in(r)) + " -> " +
                                     std::to_string(cells_out(r)));
                q->error = rg_last_error();
                continue;
            }
            const int64_t es = r.dtype == RG_F32 ? 4 : 8;
            const int64_t nsrc = r.transpose ? w->n_out : w->n_in;
            const int64_t ndst = r.transpose ? w->n_in : w->n_out;
            void *dx = nullptr, *dy = nullptr;
            int e = dev_alloc(&dx, std::max<int64_t>(r.batch * nsrc * es, 1), s);
            if (e == RG_OK) e = dev_alloc(&dy, std::max<int64_t>(r.batch * ndst * es, 1), s);
            if (e != RG_OK) {
                dev_free(dx);
                failed(e);
                break;
            }
            dev_bufs.push_back(dx);
            dev_bufs.push_back(dy);
            if (r.batch > 0 && hipMemcpyAsync(dx, r.values, r.batch * nsrc * es,
                                              hipMemcpyHostToDevice, s) != hipSuccess) {
                failed(fail(RG_ERR_HIP, "value upload failed"));
                break;
            }
            rg_apply_job j;
            j.weights = w;
            j.transpose = r.transpose;
            j.values_input = dx;
            j.values_output = dy;
            j.batch = r.batch;
            jobs[r.dtype == RG_F32 ? 1 : 0].push_back(j);
            outs.emplace_back(q, dy);
        }
    }
    if (status == RG_OK && !jobs[0].empty()) {
        int e = rg_apply_batched_f64(jobs[0].data(), (int64_t)jobs[0].size(), s);
        if (e != RG_OK) failed(e);
    }
    if (status == RG_OK && !jobs[1].empty()) {
        int e = rg_apply_batched_f32(jobs[1].data(), (int64_t)jobs[1].size(), s);
        if (e != RG_OK) failed(e);
    }
    if (status == RG_OK) {
        for (auto& o : outs) {
            const rg_serve_request& r = o.first->r;
            rg_weights* w = nullptr;
            for (auto& j : jobs[r.dtype == RG_F32 ? 1 : 0])
                if (j.values_output == o.second) w = j.weights;
            const int64_t es = r.dtype == RG_F32 ? 4 : 8;
            const int64_t ndst = r.transpose ? w->n_in : w->n_out;
            if (r.batch > 0 && hipMemcpyAsync(r.result, o.second, r.batch * ndst * es,
                                              hipMemcpyDeviceToHost, s) != hipSuccess) {
                failed(fail(RG_ERR_HIP, "result download failed"));
                break;
            }
        }
    }
    if (hipStreamSynchronize(s) != hipSuccess) failed(fail(RG_ERR_HIP, "stream failed"));
    for (void* d : dev_bufs) dev_free(d);
    for (rg_weights* w : uncached) weights_release(w);
    if (status != RG_OK) {
        // key-0 builds of a failed group were never handed out
        for (size_t k = 0; k < build_reqs.size(); k++)
            if (!build_reqs[k]->r.key && built[k] &&
                std::find(uncached.begin(), uncached.end(), built[k]) == uncached.end())
                weights_release(built[k]);
    }
    const int64_t t = now_ns();
    {
        std::lock_guard<std::mutex> lk(mu);
        for (Req* q : group) {
            if (q->r.key) {
                auto ce = cache.find(q->r.key);
                if (ce != cache.end() && ce->second.pins > 0) ce->second.pins--;
            }
            admitted -= q->footprint;
            q->done = true;
            if (!q->failed_alone) {
                q->status = status;
                q->error = error;
            }
            q->done_ns = t;
            if (q->status == RG_OK) st.completed++;
            else st.failed++;
        }
        evict_locked();
        st.cache_entries = (int64_t)cache.size();
        st.cache_bytes = cache_bytes;
    }
    cv_done.notify_all();
}

extern "C" {

int64_t rg_clock_ns(void) { return now_ns(); }

int rg_host_alloc(void** ptr, size_t bytes)
{
    if (!ptr) return fail(RG_ERR_ARGUMENT, "ptr is NULL");
    *ptr = nullptr;
    RG_HIP(hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault));
    return RG_OK;
}

int rg_host_free(void* ptr)
{
    if (ptr) RG_HIP(hipHostFree(ptr));
    return RG_OK;
}

int rg_serve_create(const rg_serve_config* config, rg_server_t* out)
{
    if (!out) return fail(RG_ERR_ARGUMENT, "out is NULL");
    *out = nullptr;
    auto* S = new rg_server();
    rg_serve_config c = {};
    if (config) c = *config;
    if (c.workers <= 0) c.workers = 3;
    if (c.workers > 16) c.workers = 16;
    if (c.max_group <= 0) c.max_group = 64;
    if (c.reserve_bytes <= 0) c.reserve_bytes = (int64_t)4 << 30;
    if (c.cache_bytes <= 0) c.cache_bytes = (int64_t)32 << 30;
    S->cfg = c;
    (void)hipGetDevice(&S->device);
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
        delete S;
        return fail(RG_ERR_HIP, "hipMemGetInfo failed");
    }
    S->budget = (int64_t)fr - c.reserve_bytes - c.cache_bytes;
    if (S->budget < ((int64_t)256 << 20)) S->budget = (int64_t)256 << 20;
    for (int k = 0; k < c.workers; k++) S->workers.emplace_back([S] { S->worker(); });
    *out = S;
    return RG_OK;
}

int rg_serve_submit(rg_server_t S, const rg_serve_request* request, uint64_t* ticket)
{
    if (!S || !request || !ticket) return fail(RG_ERR_ARGUMENT, "NULL argument");
    const rg_serve_request& r = *request;
    if (r.dtype != RG_F64 && r.dtype != RG_F32)
        return fail(RG_ERR_ARGUMENT, "dtype must be RG_F64 or RG_F32");
    if (r.nxi < 2 || r.nyi < 2 || r.nxo < 2 || r.nyo < 2)
        return fail(RG_ERR_ARGUMENT, "each grid needs at least 2 vertices along both axes");
    if (r.batch < 0 || (r.batch > 0 && (!r.values || !r.result)))
        return fail(RG_ERR_ARGUMENT, "values or result is NULL");
    auto q = std::make_unique<Req>();
    q->r = r;
    if (!q->r.arrival_ns) q->r.arrival_ns = now_ns();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->stop) return fail(RG_ERR_ARGUMENT, "the server is stopping");
        q->ticket = S->next_ticket++;
        *ticket = q->ticket;
        S->queue.push_back(q.get());
        S->reqs.emplace(q->ticket, std::move(q));
        S->st.submitted++;
    }
    S->cv_work.notify_one();
    return RG_OK;
}

int rg_serve_wait(rg_server_t S, uint64_t ticket, int64_t* done_ns)
{
    if (!S) return fail(RG_ERR_ARGUMENT, "server is NULL");
    std::unique_ptr<Req> q;
    {
        std::unique_lock<std::mutex> lk(S->mu);
        auto it = S->reqs.find(ticket);
        if (it == S->reqs.end()) return fail(RG_ERR_ARGUMENT, "unknown ticket");
        Req* p = it->second.get();
        S->cv_done.wait(lk, [&] { return p->done; });
        q = std::move(it->second);
        S->reqs.erase(it);
    }
    if (done_ns) *done_ns = q->done_ns;
    if (q->status != RG_OK) return fail(q->status, q->error);
    return RG_OK;
}

int rg_serve_stats_get(rg_server_t S, rg_serve_stats* stats)
{
    if (!S || !stats) return fail(RG_ERR_ARGUMENT, "NULL argument");
    std::lock_guard<std::mutex> lk(S->mu);
    *stats = S->st;
    return RG_OK;
}

int rg_serve_destroy(rg_server_t S)
{
    if (!S) return RG_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    for (auto& t : S->workers) t.join();
    for (auto& kv : S->cache)
        if (kv.second.w) weights_release(kv.second.w);
    delete S;
    return RG_OK;
}

}  // extern "C"
