// ppa_api.cpp -- the C ABI (include/ppa.h) over the gfx950 kernels.
//
// Replaces the cuPPA class (src/PA/cuPPA/cuPPA.cuh:84-105): the constructor's
// (N, M, p) become a ppa_graph_desc, the per-call cudaMalloc/cudaFree of the
// graph (cuPPA_1.cu:54-60,247-250) become caller-owned buffers, the stage
// loop (cuPPA_1.cu:112-226) becomes one persistent launch, and the
// printf-and-continue error handling (cuPPA_1.cu:42-46,106-110) becomes
// status codes.
#include <algorithm>
#include <cstring>
#include <new>

#include "ppa_degree.hpp"
#include "ppa_internal.hpp"

namespace ppa {

bool tuning_valid(const ppa_tuning &t)
{
    if (t.block_threads && (t.block_threads % 64 != 0 || t.block_threads > 256)) return false;
    if (t.chunk_nodes && (t.chunk_nodes % 64 != 0 || t.chunk_nodes > (1u << 20))) return false;
    if (t.sleep_cap > 64) return false;
    return true;
}

// d > kNarrowMaxD: 64-thread workgroups of wide_geometry(d).streams windows
// each, as many as the CUs hold (the LDS limits that at large d), but no more
// than the graph has claims of wide_geometry(d).claim consecutive nodes.
static int resolve_wide_shape(const ppa_ctx *ctx, uint32_t d, uint64_t work, LaunchShape *out)
{
    const ppa_tuning &tu = ctx->tuning;
    int per_cu = 0;
    PPA_HIP(generate_occupancy(d, kWideThreads, &per_cu));
    if (per_cu < 1) per_cu = 1;
    if (tu.blocks_per_cu && (int)tu.blocks_per_cu < per_cu) per_cu = (int)tu.blocks_per_cu;
    const uint32_t per_claim = wide_geometry(d).claim;
    const uint64_t claims = (work + per_claim - 1) / per_claim;
    uint64_t g = std::min<uint64_t>((uint64_t)per_cu * (uint64_t)ctx->num_cus, claims);
    if (tu.max_blocks && g > tu.max_blocks) g = tu.max_blocks;
    if (g < 1) g = 1;
    out->blocks = (uint32_t)g;
    out->threads = kWideThreads;
    out->chunk = per_claim;
    out->early_nodes = 0;
    const uint32_t wd = tu.watchdog_ms ? tu.watchdog_ms : kDefaultWatchdogMs;
    out->watchdog_ticks = (uint64_t)wd * (uint64_t)ctx->wall_clock_khz;
    out->sleep_cap = tu.sleep_cap ? tu.sleep_cap : (uint32_t)std::min<uint64_t>(32, std::max<uint64_t>(1, g / 256));
    out->wide_mask = 0;
    return PPA_OK;
}

int resolve_shape(const ppa_ctx *ctx, uint32_t d, uint64_t work, LaunchShape *out, bool cap_single)
{
    const ppa_tuning &tu = ctx->tuning;
    if (!tuning_valid(tu)) return PPA_ERR_INVALID;
    if (d > kNarrowMaxD) return resolve_wide_shape(ctx, d, work, out);
    const uint32_t th = tu.block_threads ? tu.block_threads : kDefaultBlock;
    const uint32_t ch = tu.chunk_nodes ? tu.chunk_nodes : kDefaultChunk;
    int per_cu = 0;
    PPA_HIP(generate_occupancy(d, th, &per_cu));
    if (per_cu < 1) per_cu = 1;
    if (tu.blocks_per_cu && (int)tu.blocks_per_cu < per_cu) per_cu = (int)tu.blocks_per_cu;
    // Single graphs: 16 wavefronts per CU issue the copy reads as fast as the
    // occupancy limit's 32 (d <= 8) or 20 (d = 16) and keep fewer of them
    // re-reading still-empty sources. Measured: n = 2^30, d = 8, p = 0.2 in
    // 174.2 ms against 178.3 ms; n = 2^28, d = 8, p = 0.5 in 29.4 against
    // 31.1; BA at n = 2^30, d = 8 in 116.3 against 118.6; n = 2^28, d = 2 in
    // 22.4 against 23.4. 12 per CU lost 5 % at n = 2^27, d = 16, p = 0.9.
    if (d && cap_single && !tu.blocks_per_cu) per_cu = std::min<int>(per_cu, std::max<int>(1, (int)(kSingleWavesPerCu * 64 / th)));
    uint64_t g = (uint64_t)per_cu * (uint64_t)ctx->num_cus;
    // Single graphs: one lane per 2*d generated nodes, but within 256 and
    // 1024 wavefronts unless one lane per 8*d nodes needs more. Dependency
    // chains bound a small graph's run time (a node's copy source is uniform
    // over the nodes before it): up to ~1000 wavefronts shorten it, and
    // beyond that extra lanes only re-read still-empty slots until the graph
    // is large enough to need them for throughput. Measured: n = 1e5, d = 16,
    // p = 0.1 takes 3.0 ms at 193 wavefronts against 4.3 ms at 16; n = 1e7,
    // d = 16 takes 7.5 ms at 1221 against 8.8 ms at 4883; n = 1e6, d = 2,
    // 0.36 ms at 977 against 0.68 ms at 3906. Batches: no more wavefronts than
    // claim chunks.
    uint64_t waves_needed = 0;
    if (d) {
        const uint64_t dense = (work + 128ull * d - 1) / (128ull * d);  // a lane per 2d nodes
        const uint64_t sparse = (work + 512ull * d - 1) / (512ull * d);  // a lane per 8d nodes
        // d > 16: a floor of 1024 (5-13 % off calls of n = 1e5 .. 1e6; their
        // nodes' 17..32 sources make longer waits to overlap)
        const uint64_t floor_waves = d > 16 ? 1024 : 256;
        waves_needed = std::max<uint64_t>(floor_waves, std::min(dense, std::max<uint64_t>(1024, sparse)));
    } else {
        waves_needed = (work + ch - 1) / ch;
    }
    const uint64_t blocks_needed = (waves_needed + th / 64 - 1) / (th / 64);
    if (g > blocks_needed) g = blocks_needed;
    if (tu.max_blocks && g > tu.max_blocks) g = tu.max_blocks;
    if (g < 1) g = 1;
    out->blocks = (uint32_t)g;
    out->threads = th;
    out->chunk = ch;
    out->early_nodes = tu.early_nodes ? tu.early_nodes : kDefaultEarlyNodes;
    const uint32_t wd = tu.watchdog_ms ? tu.watchdog_ms : kDefaultWatchdogMs;
    out->watchdog_ticks = (uint64_t)wd * (uint64_t)ctx->wall_clock_khz;
    // A waiting lane re-reads its source once per loop pass. A few hundred
    // wavefronts can do so at full speed; a full GPU of them would flood
    // memory with re-reads, so its all-waiting wavefronts sleep longer.
    const uint64_t waves = g * (th / 64);
    out->sleep_cap = tu.sleep_cap ? tu.sleep_cap : (uint32_t)std::min<uint64_t>(32, std::max<uint64_t>(1, waves / 256));
    // Likewise a lane re-reads the empty sources behind its current slot every
    // 2nd pass below 1024 wavefronts, every 4th up to 4096 and every 16th
    // beyond (every 2nd instead of every 4th: 1-4 % off calls of d = 16 at
    // n = 2e4 .. 1e6; every pass cost 2 % at n = 1e6, d = 8).
    out->wide_mask = waves >= 4096 ? 15u : (waves >= 1024 ? 3u : 1u);
    return PPA_OK;
}

uint64_t single_start(const ppa_ctx *ctx, const ppa_graph_desc &desc)
{
    return prefix_start(desc.n, desc.d, coin_threshold(desc.model, desc.p_threshold), ctx->tuning.prefix_nodes, true);
}

uint64_t range_start(const ppa_ctx *ctx, const ppa_graph_desc &desc, uint64_t first, uint64_t last)
{
    const uint64_t seeds = (uint64_t)desc.d + 1;  // nodes 0..d are the fill's seed rows
    if (first > 0) return first > seeds ? first : seeds;
    const uint64_t p = single_start(ctx, desc);
    return p < last ? p : (last > seeds ? last : seeds);
}

hipError_t enqueue_single(const ppa_ctx *ctx, const ppa_graph_desc &desc, int32_t *G, int64_t *row_ptr,
                          const LaunchShape &shape, DeviceScratch *scratch, hipStream_t stream, hipEvent_t after_fill,
                          uint64_t first, uint64_t last)
{
    const uint32_t d = desc.d;
    const uint64_t n = last ? last : desc.n;  // the generator stops at node n
    const uint64_t start = range_start(ctx, desc, first, n);
    const uint64_t T = coin_threshold(desc.model, desc.p_threshold);
    hipError_t e = launch_reset_scratch(scratch, start, stream);
    if (e != hipSuccess) return e;
    if (first == 0 && (e = launch_fill(FillArgs{G, row_ptr, desc.n, d, desc.model}, stream)) != hipSuccess) return e;
    if (after_fill && (e = hipEventRecord(after_fill, stream)) != hipSuccess) return e;
    if (first == 0 && start > (uint64_t)d + 1) {
        PrefixArgs pa{};
        pa.G = G;
        pa.T = T;
        pa.key0 = (uint32_t)desc.seed;
        pa.key1 = (uint32_t)(desc.seed >> 32);
        pa.model = desc.model;
        pa.prefix = start;
        pa.watchdog_ticks = shape.watchdog_ticks;
        pa.scratch = scratch;
        if ((e = launch_prefix(d, pa, stream)) != hipSuccess) return e;
    }
    if (n > start) {
        GenArgs ga{};
        ga.G = G;
        ga.n = n;
        ga.T = T;
        ga.key0 = (uint32_t)desc.seed;
        ga.key1 = (uint32_t)(desc.seed >> 32);
        ga.chunk = shape.chunk;
        ga.early_chunk = kEarlyChunk;
        ga.early_nodes = shape.early_nodes;
        ga.watchdog_ticks = shape.watchdog_ticks;
        ga.sleep_cap = shape.sleep_cap;
        ga.wide_mask = shape.wide_mask;
        ga.scratch = scratch;
        if ((e = launch_generate(d, ga, shape.blocks, shape.threads, stream)) != hipSuccess) return e;
    }
    return hipSuccess;
}

uint64_t held_bytes(const ppa_ctx *ctx)
{
    return (uint64_t)ctx->work_bytes + ctx->degree_scratch_len * sizeof(uint32_t);
}

int ensure_work(ppa_ctx *ctx, size_t need, hipStream_t stream)
{
    if (ctx->work_bytes >= need) return PPA_OK;
    if (ctx->scratch_budget && held_bytes(ctx) - ctx->work_bytes + need > ctx->scratch_budget)
        return PPA_ERR_BUDGET;
    PPA_HIP(hipStreamSynchronize(stream));
    if (ctx->work) PPA_HIP(hipFree(ctx->work));
    ctx->work = nullptr;
    ctx->work_bytes = 0;
    if (hipMalloc(&ctx->work, need) != hipSuccess) {
        ctx->work = nullptr;
        (void)hipGetLastError();
        return PPA_ERR_NOMEM;
    }
    ctx->work_bytes = need;
    return PPA_OK;
}

// Waits for the pending launch through its own end event (ev[2], recorded on
// the launch's stream after its last kernel), whatever stream the caller
// names, and reads its status and counters. The context is free again on
// every return path.
static int collect_launch(ppa_ctx *ctx)
{
    ctx->pending = false;
    PPA_HIP(hipEventSynchronize(ctx->ev[2]));
    DeviceScratch host{};
    PPA_HIP(hipMemcpy(&host, ctx->scratch, sizeof(host), hipMemcpyDeviceToHost));
    ppa_stats &s = ctx->last;
    s.nodes = host.stats[kStatNodes];
    s.slots = host.stats[kStatSlots];
    s.draws = host.stats[kStatDraws];
    s.retries = host.stats[kStatRetries];
    s.prefetch_misses = host.stats[kStatPrefetchMisses];
    s.polls = host.stats[kStatPolls];
    s.stalls = host.stats[kStatStalls];
    s.claims = host.stats[kStatClaims];
    s.sleeps = host.stats[kStatSleeps];
    s.staged_rows = host.stats[kStatStagedRows];
    s.passes = host.stats[kStatPasses];
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ctx->ev[1], ctx->ev[2]) == hipSuccess) s.kernel_ms = ms;
    if (hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[2]) == hipSuccess) s.total_ms = ms;
    if (host.status != 0) return host.status;
    // every generated node must have been finished by exactly one lane
    if (s.nodes != ctx->expect_nodes) return PPA_ERR_INTERNAL;
    return PPA_OK;
}

int finish_launch(ppa_ctx *ctx)
{
    const int rc = collect_launch(ctx);
    // rows before a failed range's end may be incomplete: no range continues it
    if (rc != PPA_OK) ctx->range_next = 0;
    return rc;
}

}  // namespace ppa

using namespace ppa;

extern "C" {

int ppa_ctx_create(ppa_ctx **out, int device, uint64_t scratch_budget_bytes)
{
    if (!out) return PPA_ERR_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return PPA_ERR_DEVICE;
    if (device < 0 || device >= count) return PPA_ERR_INVALID;
    DeviceGuard g(device);
    if (!g.ok()) return PPA_ERR_DEVICE;
    ppa_ctx *c = new (std::nothrow) ppa_ctx;
    if (!c) return PPA_ERR_NOMEM;
    c->device = device;
    c->scratch_budget = scratch_budget_bytes;
    if (hipDeviceGetAttribute(&c->num_cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        hipDeviceGetAttribute(&c->wall_clock_khz, hipDeviceAttributeWallClockRate, device) != hipSuccess) {
        delete c;
        return PPA_ERR_DEVICE;
    }
    if (c->wall_clock_khz <= 0) c->wall_clock_khz = 100000;
    if (hipMalloc(&c->scratch, sizeof(DeviceScratch)) != hipSuccess) {
        delete c;
        return PPA_ERR_NOMEM;
    }
    for (auto &e : c->ev) {
        if (hipEventCreate(&e) != hipSuccess) {
            ppa_ctx_destroy(c);
            return PPA_ERR_DEVICE;
        }
    }
    *out = c;
    return PPA_OK;
}

int ppa_ctx_destroy(ppa_ctx *ctx)
{
    if (!ctx) return PPA_ERR_INVALID;
    DeviceGuard g(ctx->device);
    if (ctx->pending) (void)hipEventSynchronize(ctx->ev[2]);  // the launch still uses the scratch
    for (auto &e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->scratch) (void)hipFree(ctx->scratch);
    if (ctx->degree_scratch) (void)hipFree(ctx->degree_scratch);
    if (ctx->work) (void)hipFree(ctx->work);
    delete ctx;
    return PPA_OK;
}

int ppa_ctx_trim(ppa_ctx *ctx)
{
    if (!ctx) return PPA_ERR_INVALID;
    if (ctx->pending) return PPA_ERR_AGAIN;
    DeviceGuard g(ctx->device);
    if (!g.ok()) return PPA_ERR_DEVICE;
    if (!ctx->work && !ctx->degree_scratch) return PPA_OK;
    PPA_HIP(hipDeviceSynchronize());  // earlier calls' kernels may still read the work area
    if (ctx->work) PPA_HIP(hipFree(ctx->work));
    ctx->work = nullptr;
    ctx->work_bytes = 0;
    if (ctx->degree_scratch) PPA_HIP(hipFree(ctx->degree_scratch));
    ctx->degree_scratch = nullptr;
    ctx->degree_scratch_len = 0;
    return PPA_OK;
}

int ppa_ctx_memory(ppa_ctx *ctx, uint64_t *held, uint64_t *budget)
{
    if (!ctx) return PPA_ERR_INVALID;
    if (held) *held = held_bytes(ctx);
    if (budget) *budget = ctx->scratch_budget;
    return PPA_OK;
}

int ppa_ctx_set_tuning(ppa_ctx *ctx, const ppa_tuning *t)
{
    if (!ctx || !t || !tuning_valid(*t)) return PPA_ERR_INVALID;
    ctx->tuning = *t;
    return PPA_OK;
}

int ppa_ctx_get_tuning(ppa_ctx *ctx, ppa_tuning *out)
{
    if (!ctx || !out) return PPA_ERR_INVALID;
    *out = ctx->tuning;
    return PPA_OK;
}

int ppa_get_stats(ppa_ctx *ctx, ppa_stats *out)
{
    if (!ctx || !out) return PPA_ERR_INVALID;
    *out = ctx->last;
    return PPA_OK;
}

int ppa_generate_async(ppa_ctx *ctx, const ppa_graph_desc *desc, int32_t *d_targets, int64_t *d_row_ptr,
                       void *stream_)
{
    if (!ctx || !d_targets) return PPA_ERR_INVALID;
    int rc = ppa_check_desc(desc);
    if (rc) return rc;
    if (((uintptr_t)d_targets & 15u) != 0) return PPA_ERR_INVALID;
    if (ctx->pending) return PPA_ERR_AGAIN;  // the previous launch still owns the scratch
    DeviceGuard g(ctx->device);
    if (!g.ok()) return PPA_ERR_DEVICE;
    hipStream_t stream = (hipStream_t)stream_;
    const uint64_t n = desc->n;
    const uint32_t d = desc->d;
    LaunchShape shape;
    const uint64_t start = single_start(ctx, *desc);
    if ((rc = resolve_shape(ctx, d, n > start ? n - start : 1, &shape)) != PPA_OK) return rc;

    ctx->last = ppa_stats{};
    ctx->last.grid_blocks = shape.blocks;
    ctx->last.block_threads = shape.threads;
    PPA_HIP(hipEventRecord(ctx->ev[0], stream));
    PPA_HIP(enqueue_single(ctx, *desc, d_targets, d_row_ptr, shape, ctx->scratch, stream, ctx->ev[1]));
    PPA_HIP(hipEventRecord(ctx->ev[2], stream));
    ctx->pending = true;
    ctx->expect_nodes = n - d - 1;
    ctx->range_next = 0;
    return PPA_OK;
}

int ppa_generate_range(ppa_ctx *ctx, const ppa_graph_desc *desc, int32_t *d_targets, int64_t *d_row_ptr,
                       uint64_t first, uint64_t last, void *stream_)
{
    if (!ctx || !d_targets) return PPA_ERR_INVALID;
    int rc = ppa_check_desc(desc);
    if (rc) return rc;
    if (((uintptr_t)d_targets & 15u) != 0 || last <= first || last > desc->n) return PPA_ERR_INVALID;
    if (ctx->pending) return PPA_ERR_AGAIN;
    // a later range continues the graph of the previous one, where it ended
    if (first > 0 && (first != ctx->range_next || d_targets != ctx->range_targets ||
                      std::memcmp(desc, &ctx->range_desc, sizeof(*desc)) != 0))
        return PPA_ERR_INVALID;
    DeviceGuard g(ctx->device);
    if (!g.ok()) return PPA_ERR_DEVICE;
    hipStream_t stream = (hipStream_t)stream_;
    const uint32_t d = desc->d;
    const uint64_t start = range_start(ctx, *desc, first, last);
    LaunchShape shape;
    if ((rc = resolve_shape(ctx, d, last > start ? last - start : 1, &shape)) != PPA_OK) return rc;

    ctx->last = ppa_stats{};
    ctx->last.grid_blocks = shape.blocks;
    ctx->last.block_threads = shape.threads;
    PPA_HIP(hipEventRecord(ctx->ev[0], stream));
    PPA_HIP(enqueue_single(ctx, *desc, d_targets, d_row_ptr, shape, ctx->scratch, stream, ctx->ev[1], first, last));
    PPA_HIP(hipEventRecord(ctx->ev[2], stream));
    ctx->pending = true;
    const uint64_t lo = first > (uint64_t)d + 1 ? first : (uint64_t)d + 1;
    ctx->expect_nodes = last > lo ? last - lo : 0;
    ctx->range_desc = *desc;
    ctx->range_targets = d_targets;
    ctx->range_next = last == desc->n ? 0 : last;
    return PPA_OK;
}

int ppa_ctx_sync(ppa_ctx *ctx, void *stream)
{
    if (!ctx) return PPA_ERR_INVALID;
    if (!ctx->pending) return PPA_OK;
    DeviceGuard g(ctx->device);
    if (!g.ok()) return PPA_ERR_DEVICE;
    (void)stream;  // the launch's own end event is waited for
    return finish_launch(ctx);
}

int ppa_generate(ppa_ctx *ctx, const ppa_graph_desc *desc, int32_t *d_targets, int64_t *d_row_ptr,
                 void *stream)
{
    int rc = ppa_generate_async(ctx, desc, d_targets, d_row_ptr, stream);
    if (rc) return rc;
    return ppa_ctx_sync(ctx, stream);
}

int ppa_degree_histogram(ppa_ctx *ctx, const ppa_graph_desc *desc, const int32_t *d_targets,
                         uint32_t *d_degree, uint64_t *d_hist, uint64_t hist_len, uint64_t *d_max_degree,
                         void *stream_)
{
    if (!ctx || !d_targets) return PPA_ERR_INVALID;
    int rc = ppa_check_desc(desc);
    if (rc) return rc;
    if (d_hist == nullptr && hist_len != 0) return PPA_ERR_INVALID;
    if (ctx->pending) return PPA_ERR_AGAIN;
    DeviceGuard g(ctx->device);
    if (!g.ok()) return PPA_ERR_DEVICE;
    hipStream_t stream = (hipStream_t)stream_;
    uint32_t *deg = d_degree;
    if (!deg) {
        if (ctx->degree_scratch_len < desc->n) {
            if (ctx->scratch_budget && ctx->work_bytes + desc->n * sizeof(uint32_t) > ctx->scratch_budget)
                return PPA_ERR_BUDGET;
            PPA_HIP(hipStreamSynchronize(stream));  // an earlier call may still use the old one
            if (ctx->degree_scratch) PPA_HIP(hipFree(ctx->degree_scratch));
            ctx->degree_scratch = nullptr;
            ctx->degree_scratch_len = 0;
            if (hipMalloc(&ctx->degree_scratch, desc->n * sizeof(uint32_t)) != hipSuccess) return PPA_ERR_NOMEM;
            ctx->degree_scratch_len = desc->n;
        }
        deg = ctx->degree_scratch;
    }
    PPA_HIP(launch_reset_scratch(ctx->scratch, 0, stream));
    // Partitioned pass (two int32 copies of the slots plus small tables, kept
    // for the next call); one atomic per slot only if they cannot be had.
    if (ensure_work(ctx, degree_work_bytes(desc->n, desc->d, (uint32_t)ctx->num_cus), stream) == PPA_OK) {
        DegreeWork w;
        degree_work_layout(ctx->work, desc->n, desc->d, (uint32_t)ctx->num_cus, &w);
        w.status = &ctx->scratch->status;
        PPA_HIP(launch_degree_partitioned(d_targets, desc->n, desc->d, deg, w, stream));
    } else {
        PPA_HIP(launch_degree_atomic(d_targets, desc->n, desc->d, deg, &ctx->scratch->status, stream));
    }
    PPA_HIP(launch_seed_degree(deg, desc->d, desc->model, stream));
    PPA_HIP(launch_degree_histogram(deg, desc->n, d_hist, hist_len, ctx->scratch, stream));
    if (d_max_degree)
        PPA_HIP(hipMemcpyAsync(d_max_degree, &ctx->scratch->max_degree, sizeof(uint64_t),
                               hipMemcpyDeviceToDevice, stream));
    PPA_HIP(hipStreamSynchronize(stream));
    int32_t status = 0;  // PPA_ERR_RANGE from a degree that passed 2^32 - 1
    PPA_HIP(hipMemcpy(&status, &ctx->scratch->status, sizeof(status), hipMemcpyDeviceToHost));
    return status;
}

int ppa_histogram_of_degrees(ppa_ctx *ctx, uint64_t n, const uint32_t *d_degree, uint64_t *d_hist,
                             uint64_t hist_len, uint64_t *d_max_degree, void *stream_)
{
    if (!ctx || !d_degree || n == 0 || n > PPA_MAX_N) return PPA_ERR_INVALID;
    if (d_hist == nullptr && hist_len != 0) return PPA_ERR_INVALID;
    if (ctx->pending) return PPA_ERR_AGAIN;
    DeviceGuard g(ctx->device);
    if (!g.ok()) return PPA_ERR_DEVICE;
    hipStream_t stream = (hipStream_t)stream_;
    PPA_HIP(launch_reset_scratch(ctx->scratch, 0, stream));
    PPA_HIP(launch_degree_histogram(d_degree, n, d_hist, hist_len, ctx->scratch, stream));
    if (d_max_degree)
        PPA_HIP(hipMemcpyAsync(d_max_degree, &ctx->scratch->max_degree, sizeof(uint64_t),
                               hipMemcpyDeviceToDevice, stream));
    PPA_HIP(hipStreamSynchronize(stream));
    return PPA_OK;
}

int ppa_in_csr(ppa_ctx *ctx, const ppa_graph_desc *desc, const int32_t *d_targets, int64_t *d_in_ptr,
               int32_t *d_in_src, void *stream_)
{
    if (!ctx || !d_targets || !d_in_ptr || !d_in_src) return PPA_ERR_INVALID;
    int rc = ppa_check_desc(desc);
    if (rc) return rc;
    if (ctx->pending) return PPA_ERR_AGAIN;
    DeviceGuard g(ctx->device);
    if (!g.ok()) return PPA_ERR_DEVICE;
    hipStream_t stream = (hipStream_t)stream_;
    if ((rc = ensure_work(ctx, csr_work_bytes(desc->n, desc->d, (uint32_t)ctx->num_cus), stream)) != PPA_OK)
        return rc;
    PPA_HIP(launch_in_csr(d_targets, desc->n, desc->d, d_in_ptr, d_in_src, ctx->work, (uint32_t)ctx->num_cus,
                          stream));
    PPA_HIP(hipStreamSynchronize(stream));
    return PPA_OK;
}

int ppa_undirected_csr(ppa_ctx *ctx, const ppa_graph_desc *desc, const int32_t *d_targets, int64_t *d_ptr,
                       int32_t *d_adj, void *stream_)
{
    if (!ctx || !d_targets || !d_ptr || !d_adj) return PPA_ERR_INVALID;
    int rc = ppa_check_desc(desc);
    if (rc) return rc;
    if (ctx->pending) return PPA_ERR_AGAIN;
    DeviceGuard g(ctx->device);
    if (!g.ok()) return PPA_ERR_DEVICE;
    if (desc->model != PPA_MODEL_COPY) return PPA_ERR_INVALID;
    hipStream_t stream = (hipStream_t)stream_;
    const uint64_t n = desc->n, d = desc->d;
    // the in-edge CSR goes to the back of the work area, after the sort's buffers
    const size_t sort_bytes = (csr_work_bytes(n, (uint32_t)d, (uint32_t)ctx->num_cus) + 255) / 256 * 256;
    const size_t ptr_bytes = ((n + 1) * 8 + 255) / 256 * 256;
    if ((rc = ensure_work(ctx, sort_bytes + ptr_bytes + n * d * 4, stream)) != PPA_OK) return rc;
    int64_t *in_ptr = (int64_t *)((char *)ctx->work + sort_bytes);
    int32_t *in_src = (int32_t *)((char *)ctx->work + sort_bytes + ptr_bytes);
    const int32_t *keys = nullptr;
    PPA_HIP(launch_in_csr(d_targets, n, (uint32_t)d, in_ptr, in_src, ctx->work, (uint32_t)ctx->num_cus, stream,
                          &keys));
    int64_t total = 0;
    PPA_HIP(hipMemcpyAsync(&total, in_ptr + n, sizeof(total), hipMemcpyDeviceToHost, stream));
    PPA_HIP(hipStreamSynchronize(stream));
    if ((uint64_t)total != n * d) return PPA_ERR_INVALID;  // an incomplete graph has no undirected form
    PPA_HIP(launch_undirected(d_targets, n, (uint32_t)d, in_ptr, keys, in_src, d_ptr, d_adj, stream));
    PPA_HIP(hipStreamSynchronize(stream));
    return PPA_OK;
}

}  // extern "C"
