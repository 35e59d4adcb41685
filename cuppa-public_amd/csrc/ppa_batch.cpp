// ppa_batch.cpp -- batched ensembles in one captured persistent launch, and
// the autotuner.
//
// The reference generates one graph per process and tunes online by doubling
// threads and stage size while its LDS queues stay half empty
// (cuPPA_2.cu:181-198). Here a batch of graphs is one work space: every
// graph's nodes are cut into chunks, the chunks are handed out round by round
// across the graphs (chunk r of every graph that has one, then chunk r+1,
// ...), so each graph is claimed in ascending order -- the condition for
// progress -- while thousands of graphs advance side by side with small
// in-flight windows, instead of one small graph at a time being entirely in
// flight and waiting on itself. The reset, fill and generation kernels are
// captured once into a hipGraph that every launch replays. The autotuner
// measures launch shapes on the caller's own graphs and keeps the fastest per
// workload signature.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "ppa_internal.hpp"

struct ppa_batch {
    ppa_ctx *ctx = nullptr;
    uint64_t count = 0;
    void *d_tables = nullptr;  // packed as in ppa::BatchTables
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t expect_nodes = 0;
    ppa::LaunchShape shape;
    // Per-graph completion: the kernels raise flags[i] (mapped, coherent host
    // memory) once graph i is complete and published; ppa_batch_sync watches
    // them and keeps, per graph, the host time from ppa_batch_launch to the
    // first look that found the flag raised.
    uint32_t *flags = nullptr;
    uint32_t *d_flags = nullptr;
    std::vector<double> done_ms;
    std::chrono::steady_clock::time_point launched;
};

namespace ppa {

size_t batch_tables_bytes(uint64_t count)
{
    return count * sizeof(BatchGraph) + (2 * (count + 1) + (count + 1) + 2 * count) * sizeof(uint64_t) +
           3 * count * sizeof(uint32_t);
}

BatchGraph graph_record(const ppa_graph_desc &dsc, uint32_t prefix_cap)
{
    BatchGraph gr;
    std::memset(&gr, 0, sizeof(gr));
    gr.n = dsc.n;
    gr.T = coin_threshold(dsc.model, dsc.p_threshold);
    gr.key0 = (uint32_t)dsc.seed;
    gr.key1 = (uint32_t)(dsc.seed >> 32);
    gr.d = dsc.d;
    gr.model = dsc.model;
    gr.inv_d = 1.0 / (double)dsc.d;
    gr.start = prefix_start(dsc.n, dsc.d, gr.T, prefix_cap, false);
    return gr;
}

namespace {

struct TableView {
    BatchGraph *graphs;
    uint64_t *slot, *row, *seg_base, *seg_round, *seg_active;
    uint32_t *order;
    uint32_t *done;  // finished nodes per graph
    uint32_t *pidx;  // graphs with an LDS-resident prefix
};

TableView view_of(char *base, uint64_t count, uint64_t nseg)
{
    TableView v;
    v.graphs = (BatchGraph *)base;
    v.slot = (uint64_t *)(base + count * sizeof(BatchGraph));
    v.row = v.slot + (count + 1);
    v.seg_base = v.row + (count + 1);
    v.seg_round = v.seg_base + (nseg + 1);
    v.seg_active = v.seg_round + nseg;
    v.order = (uint32_t *)(v.seg_active + nseg);
    v.done = v.order + count;
    v.pidx = v.done + count;
    return v;
}

}  // namespace

void build_batch_tables(const std::vector<BatchGraph> &graphs, uint64_t chunk, BatchTables *out)
{
    const uint64_t count = graphs.size();
    std::vector<uint64_t> chunks(count);
    std::vector<uint32_t> order(count);
    out->work = 0;
    out->any_row_ptr = false;
    std::vector<uint32_t> pidx;
    out->nodes = 0;
    for (uint64_t i = 0; i < count; ++i) {
        const uint64_t w = graphs[i].n - graphs[i].start;
        out->work += w;
        out->nodes += graphs[i].n - graphs[i].d - 1;
        if (graphs[i].start > (uint64_t)graphs[i].d + 1) pidx.push_back((uint32_t)i);
        chunks[i] = (w + chunk - 1) / chunk;
        order[i] = (uint32_t)i;
        out->any_row_ptr |= graphs[i].row_ptr != nullptr;
    }
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return chunks[x] > chunks[y]; });
    // segments: rounds [v_i, v_{i+1}) over the distinct positive chunk counts
    std::vector<uint64_t> seg_base{0}, seg_round, seg_active;
    uint64_t round = 0;
    for (uint64_t i = count; i-- > 0;) {  // ascending chunk counts
        const uint64_t c = chunks[order[i]];
        if (c <= round) continue;
        const uint64_t active = i + 1;  // graphs order[0..i] have more than `round` chunks
        seg_round.push_back(round);
        seg_active.push_back(active);
        seg_base.push_back(seg_base.back() + (c - round) * active);
        round = c;
    }
    const uint64_t nseg = seg_round.size();
    out->count = count;
    out->nseg = nseg;
    out->bytes.assign(count * sizeof(BatchGraph) + (2 * (count + 1) + (nseg + 1) + 2 * nseg) * sizeof(uint64_t) +
                          3 * count * sizeof(uint32_t),
                      0);
    TableView v = view_of(out->bytes.data(), count, nseg);
    std::memcpy(v.graphs, graphs.data(), count * sizeof(BatchGraph));
    v.slot[0] = v.row[0] = 0;
    for (uint64_t i = 0; i < count; ++i) {
        v.slot[i + 1] = v.slot[i] + graphs[i].n * graphs[i].d;
        v.row[i + 1] = v.row[i] + (graphs[i].n + 1);
    }
    std::memcpy(v.seg_base, seg_base.data(), (nseg + 1) * sizeof(uint64_t));
    if (nseg) {
        std::memcpy(v.seg_round, seg_round.data(), nseg * sizeof(uint64_t));
        std::memcpy(v.seg_active, seg_active.data(), nseg * sizeof(uint64_t));
    }
    std::memcpy(v.order, order.data(), count * sizeof(uint32_t));
    if (!pidx.empty()) std::memcpy(v.pidx, pidx.data(), pidx.size() * sizeof(uint32_t));
    out->nprefix = pidx.size();
    out->total_slots = v.slot[count];
    out->total_rows = out->any_row_ptr ? v.row[count] : 0;
}

hipError_t enqueue_batch(void *d_tables, const BatchTables &t, const LaunchShape &shape, DeviceScratch *scratch,
                         hipStream_t stream, uint32_t *done_flag)
{
    TableView v = view_of((char *)d_tables, t.count, t.nseg);
    hipError_t e = launch_reset_scratch(scratch, 0, stream);
    if (e != hipSuccess) return e;
    BatchFillArgs fa{v.graphs, v.slot, v.row, t.count, v.done};
    if ((e = launch_fill_batch(fa, t.total_slots, t.total_rows, stream)) != hipSuccess) return e;
    if (t.nprefix != 0) {
        PrefixArgs pa{};
        pa.graphs = v.graphs;
        pa.pidx = v.pidx;
        pa.watchdog_ticks = shape.watchdog_ticks;
        pa.scratch = scratch;
        pa.done_flag = done_flag;
        if ((e = launch_prefix_batch(pa, (uint32_t)t.nprefix, stream)) != hipSuccess) return e;
    }
    if (t.work == 0) return hipSuccess;
    GenArgs ga{};
    ga.graphs = v.graphs;
    ga.order = v.order;
    ga.seg_base = v.seg_base;
    ga.seg_round = v.seg_round;
    ga.seg_active = v.seg_active;
    ga.nseg = t.nseg;
    ga.chunk = shape.chunk;
    ga.early_chunk = shape.chunk;
    // Round-robin claims keep a few chunks of each graph in flight, so only a
    // graph's first rounds are dense enough to need slot-by-slot publication.
    ga.early_nodes = std::min<uint64_t>(shape.early_nodes, 16ull * shape.chunk);
    ga.watchdog_ticks = shape.watchdog_ticks;
    ga.sleep_cap = shape.sleep_cap;
    ga.wide_mask = shape.wide_mask;
    ga.scratch = scratch;
    ga.done_nodes = v.done;
    ga.done_flag = done_flag;
    return launch_generate_batch(ga, shape.blocks, shape.threads, stream);
}

}  // namespace ppa

namespace {

using namespace ppa;

// Graph regions [off, off + len) must not overlap.
bool disjoint(std::vector<std::pair<uint64_t, uint64_t>> r)
{
    std::sort(r.begin(), r.end());
    for (size_t i = 1; i < r.size(); ++i)
        if (r[i].first < r[i - 1].first + r[i - 1].second) return false;
    return true;
}

void destroy(ppa_batch *b)
{
    if (!b) return;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->d_tables) (void)hipFree(b->d_tables);
    if (b->flags) (void)hipHostFree(b->flags);
    delete b;
}

}  // namespace

extern "C" {

int ppa_batch_create(ppa_ctx *ctx, const ppa_graph_desc *descs, uint64_t count, int32_t *d_targets,
                     int64_t *d_row_ptr, ppa_batch **out)
{
    if (!ctx || !descs || !count || !d_targets || !out) return PPA_ERR_INVALID;
    *out = nullptr;
    std::vector<std::pair<uint64_t, uint64_t>> treg, rreg;
    std::vector<BatchGraph> graphs(count);
    for (uint64_t i = 0; i < count; ++i) {
        int rc = ppa_check_desc(&descs[i]);
        if (rc) return rc;
        if (descs[i].d > kBatchDMax) return PPA_ERR_INVALID;  // the batch generator holds d <= 16
        treg.emplace_back(descs[i].targets_offset, descs[i].n * descs[i].d);
        rreg.emplace_back(descs[i].row_ptr_offset, descs[i].n + 1);
        graphs[i] = graph_record(descs[i], ctx->tuning.prefix_nodes);
        graphs[i].G = d_targets + descs[i].targets_offset;
        graphs[i].row_ptr = d_row_ptr ? d_row_ptr + descs[i].row_ptr_offset : nullptr;
        graphs[i].targets_offset = descs[i].targets_offset;
    }
    if (!disjoint(treg) || (d_row_ptr && !disjoint(rreg))) return PPA_ERR_INVALID;
    DeviceGuard g(ctx->device);
    if (!g.ok()) return PPA_ERR_DEVICE;

    ppa_batch *b = new (std::nothrow) ppa_batch;
    if (!b) return PPA_ERR_NOMEM;
    b->ctx = ctx;
    b->count = count;
    uint64_t work = 0;
    for (const auto &gr : graphs) work += gr.n - gr.start;
    int rc = resolve_shape(ctx, 0, work, &b->shape);
    if (rc) {
        destroy(b);
        return rc;
    }
    BatchTables tables;
    build_batch_tables(graphs, b->shape.chunk, &tables);
    b->expect_nodes = tables.nodes;
    if (hipMalloc(&b->d_tables, tables.bytes.size()) != hipSuccess) {
        destroy(b);
        return PPA_ERR_NOMEM;
    }
    if (hipMemcpy(b->d_tables, tables.bytes.data(), tables.bytes.size(), hipMemcpyHostToDevice) != hipSuccess) {
        destroy(b);
        return PPA_ERR_DEVICE;
    }
    if (hipHostMalloc(&b->flags, count * sizeof(uint32_t), hipHostMallocMapped | hipHostMallocCoherent) !=
            hipSuccess ||
        hipHostGetDevicePointer((void **)&b->d_flags, b->flags, 0) != hipSuccess) {
        destroy(b);
        return PPA_ERR_NOMEM;
    }
    std::memset(b->flags, 0, count * sizeof(uint32_t));
    b->done_ms.assign(count, -1.0);
    // Capture reset + fill + generate once; launches replay the graph.
    hipStream_t cap = nullptr;
    if (hipStreamCreateWithFlags(&cap, hipStreamNonBlocking) != hipSuccess) {
        destroy(b);
        return PPA_ERR_DEVICE;
    }
    bool ok = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal) == hipSuccess;
    if (ok) {
        ok = enqueue_batch(b->d_tables, tables, b->shape, ctx->scratch, cap, b->d_flags) == hipSuccess;
        hipGraph_t graph = nullptr;
        const bool ended = hipStreamEndCapture(cap, &graph) == hipSuccess;
        ok = ok && ended;
        b->graph = graph;
    }
    (void)hipStreamDestroy(cap);
    if (!ok || hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) != hipSuccess) {
        destroy(b);
        return PPA_ERR_DEVICE;
    }
    *out = b;
    return PPA_OK;
}

int ppa_batch_launch(ppa_batch *b, void *stream_)
{
    if (!b) return PPA_ERR_INVALID;
    ppa_ctx *ctx = b->ctx;
    // the captured graph resets the context's scratch: never under a pending launch
    if (ctx->pending) return PPA_ERR_AGAIN;
    DeviceGuard g(ctx->device);
    if (!g.ok()) return PPA_ERR_DEVICE;
    hipStream_t stream = (hipStream_t)stream_;
    ctx->last = ppa_stats{};
    ctx->last.grid_blocks = b->shape.blocks;
    ctx->last.block_threads = b->shape.threads;
    // no launch of this batch is in flight (the context is not pending), so
    // its flags are free to clear
    std::memset(b->flags, 0, b->count * sizeof(uint32_t));
    std::fill(b->done_ms.begin(), b->done_ms.end(), -1.0);
    b->launched = std::chrono::steady_clock::now();
    PPA_HIP(hipEventRecord(ctx->ev[0], stream));
    PPA_HIP(hipEventRecord(ctx->ev[1], stream));
    PPA_HIP(hipGraphLaunch(b->exec, stream));
    PPA_HIP(hipEventRecord(ctx->ev[2], stream));
    ctx->pending = true;
    ctx->expect_nodes = b->expect_nodes;
    return PPA_OK;
}

int ppa_batch_sync(ppa_batch *b, void *stream)
{
    if (!b) return PPA_ERR_INVALID;
    ppa_ctx *ctx = b->ctx;
    if (!ctx->pending) return ppa_ctx_sync(ctx, stream);
    double end_ms = -1.0;
    {
        DeviceGuard g(ctx->device);
        if (!g.ok()) return PPA_ERR_DEVICE;
        // Watch the per-graph flags until the launch's end event completes.
        auto since = [&]() {
            return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - b->launched).count();
        };
        uint64_t seen = 0, scan = 0;
        for (;;) {
            const bool ended = hipEventQuery(ctx->ev[2]) != hipErrorNotReady;
            if (ended) end_ms = since();
            double now = -1.0;
            for (; scan < b->count && b->done_ms[scan] >= 0.0; ++scan) {}
            for (uint64_t i = scan; i < b->count && seen < b->count; ++i) {
                if (b->done_ms[i] >= 0.0 || __atomic_load_n(&b->flags[i], __ATOMIC_ACQUIRE) == 0) continue;
                if (now < 0.0) now = since();
                b->done_ms[i] = now;
                ++seen;
            }
            if (ended || seen == b->count) break;
        }
    }
    const int rc = ppa_ctx_sync(ctx, stream);
    // A graph of the seed clique alone has no generated node to raise its
    // flag; after a launch that succeeded, it and any graph not seen before
    // the end are complete at the end.
    if (rc == PPA_OK) {
        if (end_ms < 0.0)
            end_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - b->launched).count();
        for (double &x : b->done_ms)
            if (x < 0.0) x = end_ms;
    }
    return rc;
}

int ppa_batch_completion_ms(ppa_batch *b, double *done_ms, uint64_t count)
{
    if (!b || !done_ms || count != b->count) return PPA_ERR_INVALID;
    if (b->ctx->pending) return PPA_ERR_AGAIN;
    for (uint64_t i = 0; i < count; ++i) done_ms[i] = b->done_ms[i];
    return PPA_OK;
}

int ppa_batch_destroy(ppa_batch *b)
{
    if (!b) return PPA_ERR_INVALID;
    DeviceGuard g(b->ctx->device);
    destroy(b);
    return PPA_OK;
}

int ppa_generate_batch(ppa_ctx *ctx, const ppa_graph_desc *descs, uint64_t count, int32_t *d_targets,
                       int64_t *d_row_ptr, void *stream)
{
    ppa_batch *b = nullptr;
    int rc = ppa_batch_create(ctx, descs, count, d_targets, d_row_ptr, &b);
    if (rc) return rc;
    rc = ppa_batch_launch(b, stream);
    if (!rc) rc = ppa_batch_sync(b, stream);
    ppa_batch_destroy(b);
    return rc;
}

int ppa_autotune(ppa_ctx *ctx, const ppa_graph_desc *descs, uint64_t count, int32_t *d_targets,
                 int64_t *d_row_ptr, void *stream, ppa_tuning *best_out)
{
    if (!ctx || !descs || !count || !d_targets) return PPA_ERR_INVALID;
    if (ctx->pending) return PPA_ERR_AGAIN;  // its measurements would reset the pending launch's scratch
    // workload signature: n bucket, d, p bucket (single graph) or batch shape
    std::string sig;
    if (count == 1) {
        const ppa_graph_desc &g = descs[0];
        const int nb = (int)std::floor(std::log2((double)g.n));
        const int pb = (int)std::lround((double)coin_threshold(g.model, g.p_threshold) / 4294967296.0 * 10.0);
        sig = "g:" + std::to_string(nb) + ":" + std::to_string(g.d) + ":" + std::to_string(pb);
    } else {
        double work = 0, dsum = 0;
        for (uint64_t i = 0; i < count; ++i) {
            work += (double)descs[i].n;
            dsum += descs[i].d;
        }
        sig = "b:" + std::to_string((int)std::floor(std::log2((double)count))) + ":" +
              std::to_string((int)std::floor(std::log2(work))) + ":" + std::to_string((int)(dsum / count));
    }
    const ppa_tuning base = ctx->tuning;
    auto it = ctx->tune_cache.find(sig);
    if (it != ctx->tune_cache.end()) {
        ppa_tuning t = it->second;
        t.watchdog_ms = base.watchdog_ms;
        ctx->tuning = t;
        if (best_out) *best_out = t;
        return PPA_OK;
    }
    auto measure = [&](const ppa_tuning &t, double *ms) -> int {
        ctx->tuning = t;
        int rc;
        if (count == 1) {
            rc = ppa_generate(ctx, &descs[0], d_targets, d_row_ptr, stream);
        } else {
            ppa_batch *b = nullptr;
            rc = ppa_batch_create(ctx, descs, count, d_targets, d_row_ptr, &b);
            if (!rc) {
                rc = ppa_batch_launch(b, stream);
                if (!rc) rc = ppa_batch_sync(b, stream);
                ppa_batch_destroy(b);
            }
        }
        *ms = ctx->last.total_ms;
        return rc;
    };
    // coordinate descent over the launch shape, two passes; the starting
    // shape runs once untimed first, so that the first timed run does not
    // also pay the first touch of the output buffers (that made the starting
    // shape look 5-10 % slower than it is at 4096 graphs)
    ppa_tuning best = base;
    double best_ms = 0;
    int rc = measure(best, &best_ms);
    if (!rc) rc = measure(best, &best_ms);
    if (rc) {
        ctx->tuning = base;
        return rc;
    }
    const uint32_t threads[] = {128, 256};
    const uint32_t chunks[] = {64, 128, 256, 512, 1024};
    const uint32_t per_cu[] = {0, 4, 6};
    const uint32_t early[] = {1u << 18, 1u << 20, 1u << 22, 1u << 24};
    for (int pass = 0; pass < 2; ++pass) {
        for (int axis = 0; axis < 4; ++axis) {
            const uint32_t *vals = axis == 0 ? threads : axis == 1 ? chunks : axis == 2 ? per_cu : early;
            const int nv = axis == 0 ? 2 : axis == 1 ? 5 : axis == 2 ? 3 : (count == 1 ? 4 : 0);
            for (int v = 0; v < nv; ++v) {
                ppa_tuning t = best;
                uint32_t *field = axis == 0 ? &t.block_threads
                                : axis == 1 ? &t.chunk_nodes
                                : axis == 2 ? &t.blocks_per_cu
                                            : &t.early_nodes;
                if (*field == vals[v]) continue;
                *field = vals[v];
                double ms = 0;
                if ((rc = measure(t, &ms)) != PPA_OK) {
                    ctx->tuning = base;
                    return rc;
                }
                if (ms < best_ms) {
                    best_ms = ms;
                    best = t;
                }
            }
        }
    }
    ctx->tuning = best;
    ctx->tune_cache[sig] = best;
    if (best_out) *best_out = best;
    return PPA_OK;
}

}  // extern "C"
