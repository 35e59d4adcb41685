// ppa_server.cpp -- continuous-batching request loop on one GPU.
//
// The reference has no serving layer; this is the request-level form of
// cuPPA::generateGraph. Output memory comes from one device pool sized by the
// budget (first-fit free list, 256-byte granules); requests are admitted in
// arrival order while the pool has room and run on several streams at once,
// each launch with its own scratch. A request of solo_nodes or more runs
// alone on the single-graph generator (with 1/solo_grid_div of the grid a
// ppa_generate call would use); smaller ones are grouped into launches of the
// batch generator of ppa_batch.cpp. Either way the graphs are those of
// ppa_generate. `plain` mode is the baseline: one request per launch on one
// stream through the single-graph generator with its full grid.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <new>
#include <unordered_map>
#include <vector>

#include "ppa_internal.hpp"

namespace {

using namespace ppa;
using Clock = std::chrono::steady_clock;

constexpr uint64_t kGranule = 256;

uint64_t align_up(uint64_t x) { return (x + kGranule - 1) / kGranule * kGranule; }

struct Request {
    ppa_request req{};
    int state = PPA_REQ_QUEUED;
    int status = PPA_OK;
    uint64_t off = 0, bytes = 0;  // pool allocation (targets, then row_ptr)
    bool allocated = false;
    double submit_ms = 0, start_ms = 0, done_ms = 0;
};

struct Slot {
    hipStream_t stream = nullptr;
    DeviceScratch *scratch = nullptr;  // device
    DeviceScratch *status = nullptr;   // pinned host copy written after each launch
    void *d_tables = nullptr;
    void *h_tables = nullptr;          // pinned staging for the table upload
    size_t table_cap = 0;
    hipEvent_t done = nullptr;
    uint32_t *flags = nullptr;    // mapped, coherent host memory: graph i of a batch launch is complete
    uint32_t *d_flags = nullptr;  // the same words as the device addresses them
    bool busy = false;
    bool batch = false;
    uint64_t expect_nodes = 0;
    std::vector<uint64_t> ids;
    std::vector<char> retired;    // ids[i] already finished through its flag
};

// First-fit allocator over [0, size) in granules.
class Pool {
public:
    void reset(uint64_t size)
    {
        free_.clear();
        free_[0] = size;
        used_ = 0;
        size_ = size;
    }
    bool alloc(uint64_t bytes, uint64_t *off)
    {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                *off = it->first;
                const uint64_t rest = it->second - bytes, start = it->first + bytes;
                free_.erase(it);
                if (rest) free_[start] = rest;
                used_ += bytes;
                return true;
            }
        }
        return false;
    }
    void release(uint64_t off, uint64_t bytes)
    {
        used_ -= bytes;
        auto next = free_.lower_bound(off);
        uint64_t start = off, len = bytes;
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                start = prev->first;
                len += prev->second;
                free_.erase(prev);
            }
        }
        if (next != free_.end() && next->first == off + bytes) {
            len += next->second;
            free_.erase(next);
        }
        free_[start] = len;
    }
    uint64_t used() const { return used_; }
    uint64_t size() const { return size_; }

private:
    std::map<uint64_t, uint64_t> free_;
    uint64_t used_ = 0, size_ = 0;
};

}  // namespace

struct ppa_server {
    ppa_ctx *ctx = nullptr;
    ppa_server_config cfg{};
    char *pool_base = nullptr;
    Pool pool;
    std::unordered_map<uint64_t, Request> reqs;
    std::deque<uint64_t> queue;   // submitted, not yet admitted
    std::deque<uint64_t> ready;   // admitted (memory held), not yet launched
    std::vector<Slot> slots;
    Clock::time_point t0;
    uint64_t running = 0;

    double now_ms() const { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }
};

namespace {

uint64_t request_bytes(const ppa_request &r) { return align_up(r.n * r.d * 4) + align_up((r.n + 1) * 8); }

void destroy_server(ppa_server *s)
{
    if (!s) return;
    for (auto &sl : s->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        if (sl.scratch) (void)hipFree(sl.scratch);
        if (sl.status) (void)hipHostFree(sl.status);
        if (sl.d_tables) (void)hipFree(sl.d_tables);
        if (sl.h_tables) (void)hipHostFree(sl.h_tables);
        if (sl.flags) (void)hipHostFree(sl.flags);
    }
    if (s->pool_base) (void)hipFree(s->pool_base);
    delete s;
}

int32_t *targets_of(ppa_server *s, const Request &r) { return (int32_t *)(s->pool_base + r.off); }
int64_t *row_ptr_of(ppa_server *s, const Request &r)
{
    return (int64_t *)(s->pool_base + r.off + align_up(r.req.n * r.req.d * 4));
}

ppa_graph_desc desc_of(const ppa_request &r)
{
    ppa_graph_desc g{};
    g.n = r.n;
    g.d = r.d;
    g.model = r.model;
    g.p_threshold = r.p_threshold;
    g.seed = r.seed;
    return g;
}

// Starts one launch on slot `sl` with the requests `ids` (all allocated).
int launch(ppa_server *s, Slot &sl, const std::vector<uint64_t> &ids)
{
    std::vector<BatchGraph> graphs;
    uint64_t work = 0, expect_nodes = 0;  // the persistent generator's nodes; all nodes
    for (uint64_t id : ids) {
        Request &r = s->reqs[id];
        BatchGraph g = graph_record(desc_of(r.req), s->ctx->tuning.prefix_nodes);
        g.G = targets_of(s, r);
        g.row_ptr = row_ptr_of(s, r);
        g.targets_offset = 0;
        graphs.push_back(g);
        work += r.req.n - g.start;
        expect_nodes += r.req.n - r.req.d - 1;
    }
    // One request: the single-graph generator (compile-time d, full occupancy).
    // Several: the batch generator over all of them.
    const bool single = ids.size() == 1;
    if (single) {
        // the persistent generator's share after the LDS-resident prefix
        const uint64_t n = graphs[0].n, start = single_start(s->ctx, desc_of(s->reqs[ids[0]].req));
        work = n > start ? n - start : 1;
    }
    // The server's grid caps below are fractions of the occupancy limit, so
    // its single launches skip the single-call wavefront cap.
    LaunchShape shape;
    int rc = resolve_shape(s->ctx, single ? graphs[0].d : 0, work, &shape, false);
    if (rc) return rc;
    if (!s->cfg.plain) {
        // A persistent launch keeps its workgroups until its work is claimed.
        // Large requests take at most 3/4 of the GPU and batches of small ones
        // at most 1/4, so the two kinds overlap instead of queueing.
        LaunchShape full;
        if ((rc = resolve_shape(s->ctx, single ? graphs[0].d : 0, ~0ull >> 1, &full, false)) != PPA_OK) return rc;
        const uint32_t cap = single && graphs[0].n >= s->cfg.solo_nodes ? full.blocks * 3 / 4 : full.blocks / 4;
        shape.blocks = std::max<uint32_t>(1u, std::min(shape.blocks, cap));
        // A lone request's grid from the single-call rule is sized for its
        // latency on an idle GPU; here up to `streams` launches share it, and
        // a chain-bound grid mostly waits while holding its wave slots.
        if (single) shape.blocks = std::max<uint32_t>(1u, shape.blocks / s->cfg.solo_grid_div);
    }
    if (single) {
        const Request &r = s->reqs[ids[0]];
        PPA_HIP(enqueue_single(s->ctx, desc_of(r.req), graphs[0].G, graphs[0].row_ptr, shape, sl.scratch, sl.stream));
    } else {
        BatchTables t;
        build_batch_tables(graphs, shape.chunk, &t);
        if (t.bytes.size() > sl.table_cap) return PPA_ERR_INTERNAL;
        std::memcpy(sl.h_tables, t.bytes.data(), t.bytes.size());
        PPA_HIP(hipMemcpyAsync(sl.d_tables, sl.h_tables, t.bytes.size(), hipMemcpyHostToDevice, sl.stream));
        // the slot's previous launch has ended, so its flags are ours to clear
        std::memset(sl.flags, 0, ids.size() * sizeof(uint32_t));
        PPA_HIP(enqueue_batch(sl.d_tables, t, shape, sl.scratch, sl.stream, sl.d_flags));
    }
    sl.batch = !single;
    sl.retired.assign(ids.size(), 0);
    PPA_HIP(hipMemcpyAsync(sl.status, sl.scratch, sizeof(DeviceScratch), hipMemcpyDeviceToHost, sl.stream));
    PPA_HIP(hipEventRecord(sl.done, sl.stream));
    const double now = s->now_ms();
    for (uint64_t id : ids) {
        Request &r = s->reqs[id];
        r.state = PPA_REQ_RUNNING;
        r.start_ms = now;
    }
    sl.ids = ids;
    sl.expect_nodes = expect_nodes;
    sl.busy = true;
    s->running += ids.size();
    return PPA_OK;
}

}  // namespace

extern "C" {

int ppa_server_create(ppa_ctx *ctx, const ppa_server_config *cfg, ppa_server **out)
{
    if (!ctx || !cfg || !out || cfg->budget_bytes < kGranule) return PPA_ERR_INVALID;
    *out = nullptr;
    DeviceGuard g(ctx->device);
    if (!g.ok()) return PPA_ERR_DEVICE;
    ppa_server *s = new (std::nothrow) ppa_server;
    if (!s) return PPA_ERR_NOMEM;
    s->ctx = ctx;
    s->cfg = *cfg;
    if (!s->cfg.streams) s->cfg.streams = 16;
    if (s->cfg.streams > 16) s->cfg.streams = 16;
    if (s->cfg.plain) s->cfg.streams = 1;
    if (!s->cfg.max_batch) s->cfg.max_batch = 4096;
    if (!s->cfg.batch_nodes) s->cfg.batch_nodes = 1ull << 22;
    // Requests of 1e4 nodes and more launch alone: with the single-graph
    // kernel's grid sized to the request and the LDS-resident prefix, a lone
    // launch beats a place in a batch launch from there on (served at equal
    // p99, when this was chosen: 7.1x the plain rate, against 3.6x when
    // requests below 2^20 nodes were batched); batches collect the smallest
    // requests.
    if (!s->cfg.solo_nodes) s->cfg.solo_nodes = 10000;
    // A third of the single-call grid for lone launches: at 16 streams the
    // served rate at equal p99 rose from 5.3x to 6.2x the plain rate (a half
    // and a quarter: also 6.2x; an eighth: below 5x, the requests' own chains
    // slow down).
    if (!s->cfg.solo_grid_div) s->cfg.solo_grid_div = 3;
    const uint64_t budget = s->cfg.budget_bytes / kGranule * kGranule;
    if (hipMalloc(&s->pool_base, budget) != hipSuccess) {
        destroy_server(s);
        return PPA_ERR_NOMEM;
    }
    s->pool.reset(budget);
    s->slots.resize(s->cfg.streams);
    const size_t cap = batch_tables_bytes(s->cfg.max_batch);
    for (auto &sl : s->slots) {
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess ||
            hipMalloc(&sl.scratch, sizeof(DeviceScratch)) != hipSuccess ||
            hipHostMalloc(&sl.status, sizeof(DeviceScratch), hipHostMallocDefault) != hipSuccess ||
            hipMalloc(&sl.d_tables, cap) != hipSuccess ||
            hipHostMalloc(&sl.h_tables, cap, hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc(&sl.flags, s->cfg.max_batch * sizeof(uint32_t),
                          hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
            hipHostGetDevicePointer((void **)&sl.d_flags, sl.flags, 0) != hipSuccess) {
            destroy_server(s);
            return PPA_ERR_NOMEM;
        }
        sl.table_cap = cap;
    }
    s->t0 = Clock::now();
    *out = s;
    return PPA_OK;
}

int ppa_server_destroy(ppa_server *s)
{
    if (!s) return PPA_ERR_INVALID;
    DeviceGuard g(s->ctx->device);
    destroy_server(s);
    return PPA_OK;
}

int ppa_server_submit(ppa_server *s, const ppa_request *req)
{
    if (!s || !req) return PPA_ERR_INVALID;
    const ppa_graph_desc dsc = desc_of(*req);
    int rc = ppa_check_desc(&dsc);
    if (rc) return rc;
    if (s->reqs.count(req->id)) return PPA_ERR_INVALID;
    if (request_bytes(*req) > s->pool.size()) return PPA_ERR_BUDGET;
    Request r;
    r.req = *req;
    r.bytes = request_bytes(*req);
    r.submit_ms = s->now_ms();
    s->reqs.emplace(req->id, r);
    s->queue.push_back(req->id);
    return PPA_OK;
}

int ppa_server_step(ppa_server *s)
{
    if (!s) return PPA_ERR_INVALID;
    DeviceGuard g(s->ctx->device);
    if (!g.ok()) return PPA_ERR_DEVICE;
    int finished = 0;
    // 1. retire finished requests: graphs of a running batch launch as their
    //    flags rise (the kernel sets one only after the graph is published),
    //    then whole launches as their events complete
    for (auto &sl : s->slots) {
        if (!sl.busy) continue;
        if (sl.batch) {
            double now = -1;
            for (size_t i = 0; i < sl.ids.size(); ++i) {
                if (sl.retired[i] || __atomic_load_n(&sl.flags[i], __ATOMIC_ACQUIRE) == 0) continue;
                if (now < 0) now = s->now_ms();
                Request &r = s->reqs[sl.ids[i]];
                r.state = PPA_REQ_DONE;
                r.status = PPA_OK;
                r.done_ms = now;
                sl.retired[i] = 1;
                --s->running;
                ++finished;
            }
        }
        const hipError_t q = hipEventQuery(sl.done);
        if (q == hipErrorNotReady) continue;
        int status = q == hipSuccess ? sl.status->status : PPA_ERR_DEVICE;
        if (status == PPA_OK && sl.status->stats[kStatNodes] != sl.expect_nodes) status = PPA_ERR_INTERNAL;
        const double now = s->now_ms();
        for (size_t i = 0; i < sl.ids.size(); ++i) {
            if (sl.retired[i]) continue;  // complete before the launch ended: valid whatever the status
            Request &r = s->reqs[sl.ids[i]];
            r.state = status == PPA_OK ? PPA_REQ_DONE : PPA_REQ_FAILED;
            r.status = status;
            r.done_ms = now;
            --s->running;
            ++finished;
        }
        sl.ids.clear();
        sl.retired.clear();
        sl.busy = false;
    }
    // 2. admit in arrival order while the pool has room
    while (!s->queue.empty()) {
        Request &r = s->reqs[s->queue.front()];
        if (!s->pool.alloc(r.bytes, &r.off)) break;
        r.allocated = true;
        s->ready.push_back(s->queue.front());
        s->queue.pop_front();
    }
    // 3. start launches on free streams
    for (auto &sl : s->slots) {
        if (sl.busy || s->ready.empty()) continue;
        std::vector<uint64_t> ids;
        uint64_t nodes = 0;
        while (!s->ready.empty() && ids.size() < s->cfg.max_batch) {
            const Request &r = s->reqs[s->ready.front()];
            // (the batch generator holds d <= kBatchDMax: wider requests launch alone)
            const bool solo = s->cfg.plain || r.req.n >= s->cfg.solo_nodes || r.req.d > kBatchDMax;
            if (solo && !ids.empty()) break;
            ids.push_back(s->ready.front());
            s->ready.pop_front();
            nodes += r.req.n;
            if (solo || nodes >= s->cfg.batch_nodes) break;
        }
        int rc = launch(s, sl, ids);
        if (rc) {
            for (uint64_t id : ids) {
                Request &r = s->reqs[id];
                r.state = PPA_REQ_FAILED;
                r.status = rc;
                r.done_ms = s->now_ms();
                ++finished;
            }
            return rc;
        }
    }
    return finished;
}

int ppa_server_poll(ppa_server *s, uint64_t id, ppa_result *out)
{
    if (!s || !out) return PPA_ERR_INVALID;
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return PPA_ERR_NOT_FOUND;
    const Request &r = it->second;
    std::memset(out, 0, sizeof(*out));
    out->state = r.state;
    out->status = r.status;
    out->n = r.req.n;
    out->d = r.req.d;
    out->submit_ms = r.submit_ms;
    out->start_ms = r.start_ms;
    out->done_ms = r.done_ms;
    if (r.allocated) {
        out->targets = targets_of(s, r);
        out->row_ptr = row_ptr_of(s, r);
    }
    return PPA_OK;
}

int ppa_server_release(ppa_server *s, uint64_t id)
{
    if (!s) return PPA_ERR_INVALID;
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return PPA_ERR_NOT_FOUND;
    Request &r = it->second;
    if (r.state == PPA_REQ_QUEUED || r.state == PPA_REQ_RUNNING) return PPA_ERR_AGAIN;
    if (r.allocated) s->pool.release(r.off, r.bytes);
    s->reqs.erase(it);
    return PPA_OK;
}

int ppa_server_cancel(ppa_server *s, uint64_t id)
{
    if (!s) return PPA_ERR_INVALID;
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return PPA_ERR_NOT_FOUND;
    Request &r = it->second;
    if (r.state != PPA_REQ_QUEUED) return PPA_ERR_AGAIN;
    auto q = std::find(s->queue.begin(), s->queue.end(), id);
    if (q != s->queue.end()) {
        s->queue.erase(q);
    } else {
        auto rd = std::find(s->ready.begin(), s->ready.end(), id);
        if (rd != s->ready.end()) s->ready.erase(rd);
        if (r.allocated) s->pool.release(r.off, r.bytes);
        r.allocated = false;
    }
    r.state = PPA_REQ_CANCELLED;
    r.status = PPA_ERR_CANCELLED;
    return PPA_OK;
}

int ppa_server_usage(ppa_server *s, uint64_t *used, uint64_t *budget, uint64_t *queued, uint64_t *running)
{
    if (!s) return PPA_ERR_INVALID;
    if (used) *used = s->pool.used();
    if (budget) *budget = s->pool.size();
    if (queued) *queued = s->queue.size() + s->ready.size();
    if (running) *running = s->running;
    return PPA_OK;
}

int ppa_server_replay(ppa_server *s, const ppa_request *reqs, const double *arrival_ms, uint64_t count,
                      double *latency_ms, double *wait_ms, double *makespan_ms)
{
    if (!s || (count && (!reqs || !arrival_ms || !latency_ms))) return PPA_ERR_INVALID;
    std::unordered_map<uint64_t, uint64_t> index;
    const Clock::time_point start = Clock::now();
    const double base = s->now_ms();
    uint64_t next = 0, done = 0;
    double last = 0;
    std::vector<uint64_t> live;
    while (done < count) {
        const double now = std::chrono::duration<double, std::milli>(Clock::now() - start).count();
        while (next < count && arrival_ms[next] <= now) {
            int rc = ppa_server_submit(s, &reqs[next]);
            if (rc) return rc;
            s->reqs[reqs[next].id].submit_ms = base + arrival_ms[next];
            index[reqs[next].id] = next;
            live.push_back(reqs[next].id);
            ++next;
        }
        int fin = ppa_server_step(s);
        if (fin < 0) return fin;
        if (fin > 0) {
            std::vector<uint64_t> still;
            for (uint64_t id : live) {
                Request &r = s->reqs[id];
                if (r.state == PPA_REQ_DONE || r.state == PPA_REQ_FAILED) {
                    if (r.state == PPA_REQ_FAILED) return r.status;
                    const uint64_t i = index[id];
                    latency_ms[i] = r.done_ms - (base + arrival_ms[i]);
                    if (wait_ms) wait_ms[i] = r.start_ms - (base + arrival_ms[i]);
                    last = std::max(last, r.done_ms - base);
                    if (!s->cfg.keep_results) ppa_server_release(s, id);
                    ++done;
                } else {
                    still.push_back(id);
                }
            }
            live.swap(still);
        }
        // Kept results never free memory: a queued request that does not fit
        // while nothing runs would wait forever.
        if (s->cfg.keep_results && s->running == 0 && s->ready.empty() && !s->queue.empty())
            return PPA_ERR_BUDGET;
    }
    if (makespan_ms) *makespan_ms = last;
    return PPA_OK;
}

}  // extern "C"
