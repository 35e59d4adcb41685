// ppa_kernels.hip -- gfx950 kernels of the copy-model generator.
//
// Reference kernels replaced here (src/PA/cuPPA/):
//   initial_graph_rrp (cuPPA_RRP.cu:51-76) and the per-stage
//   initialize_vertex_structures_rrp (cuPPA_RRP.cu:78-102)  -> fill_kernel, fill_batch_kernel
//   generate_graph_rrp_by_batch (cuPPA_RRP_1.cu:28-209),
//   re_execute_copy_model_rrp (cuPPA_RRP.cu:9-49), the cuBSTVertex duplicate
//   check (cuBSTVertex.cuh:153-319) and the host stage loop
//   (cuPPA_1.cu:112-226)                                       -> generate_kernel
//   find_max_queue_size (cuPPA_RRP_1.cu:9-26)                  -> wave reductions
//                                                                 into the stats
//   printDegreeDistribution's host loops (cuPPA.cu:115-160)    -> degree kernels
//
// Design (one persistent launch instead of host stages):
//   * Every lane owns one node at a time. Wavefronts claim chunks of
//     consecutive nodes from one global counter, in ascending order, and hand
//     them to their idle lanes in ascending order. A slot only ever waits on a
//     slot of a smaller node, and the smallest unfinished node is always owned
//     by a lane whose own dependencies are all final and published, so the
//     grid always makes progress, whatever the residency of workgroups.
//   * Slots are write-once: the fill kernel sets every generated slot to -1
//     before the launch and the owner later stores the final value with an
//     agent-scope store. A plain (L2-cached) read therefore returns either -1
//     or the final value; only a -1 needs a coherent (agent-scope) re-read.
//   * On claiming a node, a lane draws attempt 0 of all d slots at once and
//     issues their d copy reads together (d-way memory-level parallelism);
//     the duplicate check then walks the slots in order from registers. Only a
//     rejected attempt (duplicate) or an empty source slot costs another
//     round trip. An empty source is kept in its slot as -2 - source, and a
//     node's empty sources are re-read together (the first one every pass,
//     the others every 2nd, 4th or, in large grids, 16th), so a node waits for its slowest source, not
//     for the sum of its sources' waits. The reference instead parks
//     placeholders -(v*M+k) in a per-thread LDS queue and spins
//     (cuPPA_RRP_1.cu:97-185).
//   * A wavefront whose lanes all wait sleeps with a doubling back-off capped
//     per launch: short for small grids, whose dependency chains set their
//     run time, and long for a full GPU, where re-reads would crowd out the
//     copy reads.
//   * Finished rows are published per loop iteration through LDS: the rows
//     finished by a wavefront within 64 consecutive nodes are transposed so
//     that each store instruction writes whole, contiguous 32-byte sectors
//     (agent-scope buffer stores). Per-lane stores of a 32-byte row would cost
//     a partial-sector write per 4 or 8 bytes.
//   * Below `early_nodes`, where almost every copy source is still being
//     generated, each slot is published as soon as it is accepted and claims
//     are one node per lane, which shortens the dependency chains.
//   * The first 1,000-2,500 nodes of a graph, where the chains are steepest,
//     come from prefix_kernel: one workgroup with the graph in LDS (see the
//     LDS-resident prefix section below); the persistent launch then claims
//     from the prefix's end.
#include "ppa_device.hpp"
#include "ppa_kernels.hpp"
#include "ppa_model.hpp"

namespace ppa {
namespace {

typedef unsigned v4u __attribute__((ext_vector_type(4)));
constexpr int kBufferFlags = 0x00020000;  // gfx9 buffer descriptor word 3 (32-bit data format)
constexpr int kSc1 = 16;                  // cache policy: agent scope (sc1)
constexpr int kNt = 2;                    // cache policy: non-temporal (nt)

// A slot whose attempt-0 copy source was still empty holds -2 - source until
// the source is final (-1 keeps meaning "no value"); sources at or beyond
// kPendingLimit are re-derived instead.
constexpr uint64_t kPendingLimit = 0x7ffffffeull;
constexpr int32_t kFarPair = 0x7fffffff;  // a copy source at or beyond kPendingLimit
constexpr int kWaitVm0 = 0x0F70;          // s_waitcnt vmcnt(0), other counters untouched
constexpr int kWaitLgkm0 = 0xC07F;        // s_waitcnt lgkmcnt(0), other counters untouched
__device__ __forceinline__ int32_t pending_code(uint64_t src) { return (int32_t)(-2 - (int64_t)src); }
__device__ __forceinline__ uint64_t pending_src(int32_t v) { return (uint64_t)(uint32_t)(-2 - v); }

__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc_of(void *base, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(base, 0, (int)bytes, kBufferFlags);
}

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t lo = __shfl_xor((uint32_t)v, off, kWave);
        const uint32_t hi = __shfl_xor((uint32_t)(v >> 32), off, kWave);
        const uint64_t o = ((uint64_t)hi << 32) | lo;
        v = o < v ? o : v;
    }
    return v;
}

__device__ __forceinline__ uint64_t wave_or_u64(uint64_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t lo = __shfl_xor((uint32_t)v, off, kWave);
        const uint32_t hi = __shfl_xor((uint32_t)(v >> 32), off, kWave);
        v |= ((uint64_t)hi << 32) | lo;
    }
    return v;
}

__device__ __forceinline__ uint64_t readlane64(uint64_t v, uint32_t lane)
{
    const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)v, (int)lane);
    const uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), (int)lane);
    return ((uint64_t)hi << 32) | lo;
}

__global__ void reset_scratch_kernel(DeviceScratch *s, uint64_t first)
{
    if (threadIdx.x == 0) {
        s->claim = first;
        s->status = 0;
        s->max_degree = 0;
    }
    if (threadIdx.x < kStatCount) s->stats[threadIdx.x] = 0;
}

// Sets generated slots to -1, writes the seed rows 0..d and row_ptr[t] = t*d.
// A one-shot grid of small workgroups, each storing one span and exiting:
// blocks [0, slot_blocks) cover the slots, 16-byte groups at a time (the host
// guarantees G is 16-byte aligned), the rest row_ptr. Non-temporal stores:
// nothing reads the span back before the generator. This wrote 6.4 TB/s in a
// probe of store patterns, against 5.8 for the persistent grid-stride loop
// it replaces (16384 workgroups).
constexpr uint32_t kFillThreads = 64;
constexpr uint32_t kFillUnroll = 4;
constexpr uint32_t kFillSpan = kFillThreads * kFillUnroll;  // 16-byte groups (or row_ptr entries) per block

__global__ __launch_bounds__(kFillThreads) void fill_kernel(FillArgs a, uint64_t slot_blocks, uint64_t blocks)
{
    typedef int v4i __attribute__((ext_vector_type(4)));
    const uint64_t slots = a.n * a.d;
    const uint64_t clique = (uint64_t)(a.d + 1) * a.d;
    const uint64_t full = slots / 4;  // whole 16-byte groups
    v4i *const G4 = reinterpret_cast<v4i *>(a.G);
    for (uint64_t blk = blockIdx.x; blk < blocks; blk += gridDim.x) {
        if (blk < slot_blocks) {
            const uint64_t base = blk * kFillSpan + threadIdx.x;
#pragma unroll
            for (uint32_t u = 0; u < kFillUnroll; ++u) {
                const uint64_t g = base + (uint64_t)u * kFillThreads;
                const uint64_t e = g * 4;
                if (g < full) {
                    if (e >= clique) {
                        __builtin_nontemporal_store((v4i){kEmpty, kEmpty, kEmpty, kEmpty}, G4 + g);
                    } else {
                        v4i w;
#pragma unroll
                        for (int i = 0; i < 4; ++i) {
                            const uint64_t q = e + i;
                            w[i] = q < clique ? seed_target(a.model, a.d, q / a.d, (uint32_t)(q % a.d)) : kEmpty;
                        }
                        G4[g] = w;
                    }
                } else if (g == full) {
                    for (uint64_t q = e; q < slots; ++q)
                        a.G[q] = q < clique ? seed_target(a.model, a.d, q / a.d, (uint32_t)(q % a.d)) : kEmpty;
                }
            }
        } else if (a.row_ptr) {
            const uint64_t base = (blk - slot_blocks) * kFillSpan + threadIdx.x;
#pragma unroll
            for (uint32_t u = 0; u < kFillUnroll; ++u) {
                const uint64_t t = base + (uint64_t)u * kFillThreads;
                if (t <= a.n) __builtin_nontemporal_store((int64_t)(t * a.d), a.row_ptr + t);
            }
        }
    }
}

// Largest i in [0, count) with base[i] <= x (base ascending, base[0] = 0).
__device__ __forceinline__ uint64_t find_segment(const uint64_t *base, uint64_t count, uint64_t x)
{
    uint64_t lo = 0, hi = count;
    while (hi - lo > 1) {
        const uint64_t mid = (lo + hi) / 2;
        if (base[mid] <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}

constexpr uint32_t kFillPerThread = 64;  // batch fill: a span of 64 KB per block and step

// The fill of every graph of a batch over one flat index space of slots and
// one of row offsets. A block locates the graph of its span once (a binary
// search over the graphs, whose latency the 64 KB span amortises: 8 KB
// spans filled at 4.4 TB/s, 64 KB ones at 6.4); a span inside one graph
// (nearly all of them: graphs hold up to 1.6e8 slots) is filled without
// per-slot lookups, and only the (d+1)*d seed slots at a graph's start
// compute a value.
__device__ __forceinline__ int32_t fill_value(const BatchGraph &gr, uint64_t local)
{
    const uint32_t clique = (gr.d + 1) * gr.d;
    if (local >= clique) return kEmpty;
    const uint32_t l = (uint32_t)local;
    return seed_target(gr.model, gr.d, l / gr.d, l % gr.d);
}

__global__ __launch_bounds__(256) void fill_batch_kernel(BatchFillArgs a, uint64_t total_slots,
                                                          uint64_t total_rows)
{
    const uint64_t per_block = (uint64_t)blockDim.x * kFillPerThread;
    if (a.done_nodes)
        for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < a.count;
             g += (uint64_t)gridDim.x * blockDim.x)
            a.done_nodes[g] = 0;
    for (uint64_t base = blockIdx.x * per_block; base < total_slots; base += gridDim.x * per_block) {
        const uint64_t end = base + per_block < total_slots ? base + per_block : total_slots;
        uint64_t g = find_segment(a.slot_base, a.count, base);
        if (a.slot_base[g + 1] >= end) {
            const BatchGraph &gr = a.graphs[g];
            int32_t *const G = gr.G;
            const uint64_t first = base - a.slot_base[g];
            for (uint32_t j = 0; j < kFillPerThread; ++j) {
                const uint64_t i = (uint64_t)j * blockDim.x + threadIdx.x;
                if (base + i >= end) break;
                G[first + i] = fill_value(gr, first + i);
            }
        } else {
            for (uint32_t j = 0; j < kFillPerThread; ++j) {
                const uint64_t s = base + (uint64_t)j * blockDim.x + threadIdx.x;
                if (s >= end) break;
                while (s >= a.slot_base[g + 1]) ++g;
                const BatchGraph &gr = a.graphs[g];
                const uint64_t local = s - a.slot_base[g];
                gr.G[local] = fill_value(gr, local);
            }
        }
    }
    for (uint64_t base = blockIdx.x * per_block; base < total_rows; base += gridDim.x * per_block) {
        const uint64_t end = base + per_block < total_rows ? base + per_block : total_rows;
        uint64_t g = find_segment(a.row_base, a.count, base);
        if (a.row_base[g + 1] >= end) {
            const BatchGraph &gr = a.graphs[g];
            if (gr.row_ptr) {
                const uint64_t first = base - a.row_base[g];
                for (uint32_t j = 0; j < kFillPerThread; ++j) {
                    const uint64_t i = (uint64_t)j * blockDim.x + threadIdx.x;
                    if (base + i >= end) break;
                    gr.row_ptr[first + i] = (int64_t)(gr.targets_offset + (first + i) * gr.d);
                }
            }
        } else {
            for (uint32_t j = 0; j < kFillPerThread; ++j) {
                const uint64_t s = base + (uint64_t)j * blockDim.x + threadIdx.x;
                if (s >= end) break;
                while (s >= a.row_base[g + 1]) ++g;
                const BatchGraph &gr = a.graphs[g];
                if (gr.row_ptr) {
                    const uint64_t t = s - a.row_base[g];
                    gr.row_ptr[t] = (int64_t)(gr.targets_offset + t * gr.d);
                }
            }
        }
    }
}

// The generator. DMAX is the out-degree (single graph, BATCH = false) or
// its upper bound (BATCH = true: every lane carries its graph's parameters).
constexpr int kFlushPasses = 3;

// Wavefronts per SIMD each out-degree is compiled for: the most whose register
// budget (512 / waves) holds the kernel without spills.
constexpr int waves_per_simd(int dmax, bool batch)
{
    return batch ? 4
                 : (dmax <= 5 || dmax == 8) ? 8
                 : (dmax <= 7 || dmax == 12) ? 6
                 : (dmax <= 11 || dmax == 16) ? 5
                 : dmax <= 16 ? 4
                 : 2;  // d = 17..32: up to 256 VGPRs, no spills
}

template <int DMAX, bool BATCH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(waves_per_simd(DMAX, BATCH))))
void generate_kernel(GenArgs a)
{
    // attempt-0 reads: non-temporal at d <= 8 (1 % faster at n = 2^30, d = 8);
    // default policy above (3-5 % faster at d = 16)
    constexpr int kPrefetchPolicy = DMAX <= 8 && !BATCH ? kNt : 0;
    __shared__ __attribute__((aligned(16))) int32_t stage_lds[4][64 * DMAX];
    __shared__ int32_t info_lds[4][64 * DMAX];
    const uint32_t lane = lane_id();
    int32_t *const stage = stage_lds[threadIdx.x >> 6];
    int32_t *const info = info_lds[threadIdx.x >> 6];
    const uint64_t start_clock = wall_clock64();

    // Wavefront-uniform claim window [wcur, wend) of node (or work) indices,
    // and for batches the window's graph.
    uint64_t wcur = 0, wend = 0, wlast = 0, fresh_base = 0;
    // Single graphs: the nodes assigned in one pass may span the end of one
    // claim window and the start of the next; fresh node i < fresh_split is
    // fresh_base + i, later ones fresh_base2 + (i - fresh_split).
    uint64_t fresh_base2 = 0;
    uint32_t fresh_split = kWave;
    bool wdone = false;
    int32_t *wG = a.G;
    uint64_t wT = a.T, wtoff = 0;
    uint32_t wk0 = a.key0, wk1 = a.key1, wd = DMAX, wgid = 0;
    double winv = 0.0;

    // Lane state.
    uint64_t t = 0;
    bool active = false, fresh = false, need_draw = false, have_src = false, early = false;
    uint32_t k = 0, r = 0;
    uint32_t skip = 0, misses = 0;  // re-read back-off after a still-empty source
    uint64_t src = 0;
    int32_t *G = a.G;
    uint64_t T = a.T;
    uint32_t k0 = a.key0, k1 = a.key1, d = DMAX, gid = 0;
    double inv_d = 0.0;
    int32_t val[DMAX];
#pragma unroll
    for (int i = 0; i < DMAX; ++i) val[i] = kEmpty;

    // Statistics: wavefront-uniform counters fed by ballots; per-lane only
    // where the amount depends on the lane's out-degree.
    // (64-bit: one wavefront of a small grid may finish 2^30 nodes and more slots)
    uint64_t s_nodes = 0, s_retries = 0, s_claims = 0, s_sleeps = 0, s_staged = 0;
    uint64_t c_draws = 0, c_slots = 0;
    uint32_t c_pmiss = 0, c_polls = 0, c_stalls = 0;
    uint32_t iter = 0, backoff = 1;
    int fail = 0;

    while (true) {
        // ---- claim work for idle lanes (wavefront-uniform control flow) ----
        bool progressed = false;
        const uint64_t idle_mask = __ballot(!active);
        if (!wdone && idle_mask != 0) {
            if (wcur >= wend) {
                const uint32_t size = (!BATCH && wlast < a.early_nodes) ? a.early_chunk : a.chunk;
                uint64_t base = 0;
                if (lane == 0) base = atomicAdd(&a.scratch->claim, (unsigned long long)size);
                base = uniform64(base);
                ++s_claims;
                wlast = base;
                if (BATCH) {
                    // global chunk index -> (segment, round, graph)
                    const uint64_t c = base / size;
                    if (c >= a.seg_base[a.nseg]) {
                        wdone = true;
                    } else {
                        const uint64_t sg = find_segment(a.seg_base, a.nseg, c);
                        const uint64_t off = c - a.seg_base[sg];
                        const uint64_t act = a.seg_active[sg];
                        const uint64_t round = a.seg_round[sg] + off / act;
                        wgid = a.order[off % act];
                        const BatchGraph &gr = a.graphs[wgid];
                        const uint64_t work = gr.n - gr.start;
                        wcur = round * size;
                        wend = wcur + size < work ? wcur + size : work;
                        wG = gr.G;
                        wT = gr.T;
                        wk0 = gr.key0;
                        wk1 = gr.key1;
                        wd = gr.d;
                        winv = gr.inv_d;
                        wtoff = gr.start;
                    }
                } else if (base >= a.n) {
                    wdone = true;
                } else {
                    wcur = base;
                    wend = base + size < a.n ? base + size : a.n;
                }
            }
            if (!wdone) {
                const uint64_t avail = wend - wcur;
                const uint32_t rank = rank_below(idle_mask);
                const uint64_t idle = (uint64_t)__popcll(idle_mask);
                fresh_base = wcur + wtoff;
                fresh_split = kWave;
                // Single graphs: a window that cannot feed every idle lane is
                // topped up with the next chunk in the same pass (claims stay
                // ascending); otherwise its last few nodes would take a pass
                // of their own with the other idle lanes doing nothing.
                uint64_t avail2 = 0;
                if (!BATCH && avail < idle) {
                    const uint32_t size = wlast < a.early_nodes ? a.early_chunk : a.chunk;
                    uint64_t base = 0;
                    if (lane == 0) base = atomicAdd(&a.scratch->claim, (unsigned long long)size);
                    base = uniform64(base);
                    ++s_claims;
                    wlast = base;
                    if (base >= a.n) {
                        wdone = true;  // the old window's remainder is assigned below
                    } else {
                        fresh_split = (uint32_t)avail;
                        fresh_base2 = base;
                        avail2 = base + size < a.n ? size : a.n - base;
                    }
                }
                if (!active && rank < avail + avail2) {
                    t = rank < avail ? wcur + rank + wtoff : fresh_base2 + (rank - avail);
                    active = true;
                    fresh = true;
                    progressed = true;
                    if (BATCH) {
                        G = wG;
                        T = wT;
                        k0 = wk0;
                        k1 = wk1;
                        d = wd;
                        inv_d = winv;
                        gid = wgid;
                    }
                    early = t < a.early_nodes;
                }
                if (avail2 != 0) {
                    wcur = fresh_base2 + (idle - avail < avail2 ? idle - avail : avail2);
                    wend = fresh_base2 + avail2;
                } else {
                    wcur += idle < avail ? idle : avail;
                }
            }
        }
        if (wdone && __ballot(active) == 0) break;

        // ---- attempt 0 of the nodes assigned this pass, over the whole wavefront ----
        // The f lanes assigned this pass hold, in rank order, the consecutive
        // nodes fresh_base + 0 .. of one claim window (single graphs: then
        // fresh_base2 + 0 .. of the next). Lane a % 64 computes pair
        // a = i*d + kk (node i, slot kk) in pass a / 64 -- ceil(f*d / 64)
        // Philox passes where per-lane draws would cost d. A copy read goes
        // straight into word a of this wavefront's LDS stage (global_load_lds:
        // no register waits on it); word a of `info` keeps ~node for a direct
        // draw, else the source pair (kFarPair when it cannot be kept as a
        // pending code). The assigned lanes then copy their info words into
        // their value registers, which frees `info` for the re-reads below:
        // both sets of reads are in flight together and the pass waits once.
        const uint64_t fresh_mask = __ballot(fresh);
        uint32_t fd = 0, first = 0;
        if (fresh_mask != 0) {
            fd = BATCH ? wd : (uint32_t)DMAX;
            const uint32_t total = (uint32_t)__popcll(fresh_mask) * fd;
            const uint32_t magic = (65536u + fd - 1u) / fd;  // exact a / fd for a < 1024, fd <= 16
#pragma unroll 1
            for (uint32_t p = 0; p * 64u < total; ++p) {
                const uint32_t ai = p * 64u + lane;
                if (ai < total) {
                    const uint32_t i = BATCH ? (ai * magic) >> 16 : ai / (uint32_t)DMAX;
                    const uint32_t kk = ai - i * fd;
                    const uint64_t tt =
                        BATCH || i < fresh_split ? fresh_base + i : fresh_base2 + (i - fresh_split);
                    const Attempt at = attempt(wk0, wk1, wT, tt, kk, 0u, tt * fd);
                    if (at.direct) {
                        info[ai] = ~(int32_t)(BATCH ? div_small(at.pair, fd, winv) : at.pair / (uint64_t)DMAX);
                    } else {
                        info[ai] = at.pair < kPendingLimit ? (int32_t)at.pair : kFarPair;
                        __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)(wG + at.pair),
                                                         (__attribute__((address_space(3))) void *)(stage + p * 64u),
                                                         4, 0, kPrefetchPolicy);
                    }
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (fresh) {
                first = rank_below(fresh_mask) * fd;
#pragma unroll
                for (int kk = 0; kk < DMAX; ++kk)
                    if (!BATCH || (uint32_t)kk < fd) val[kk] = info[first + kk];
            }
            // these LDS reads complete before any re-read below may land in info
            __builtin_amdgcn_s_waitcnt(kWaitLgkm0);
            __builtin_amdgcn_wave_barrier();
        }

        const uint32_t dd = BATCH ? d : (uint32_t)DMAX;
        const uint64_t td = t * (uint64_t)dd;
        auto node_of = [&](uint64_t pair) -> int32_t {
            return (int32_t)(BATCH ? div_small(pair, dd, inv_d) : pair / (uint64_t)DMAX);
        };

        // ---- redraws and the re-reads of still-empty sources (lanes not just assigned) ----
        // One coherent re-read per pass for the source of a redrawn slot k, and
        // for the attempt-0 sources still empty: slot k every pass (batches:
        // with a back-off), the later ones every 2nd, 4th or 16th pass. A lane's
        // re-reads land in rows kk of the LDS info area (a redrawn slot k is
        // never also pending, so its row is free); acceptance stays in slot
        // order, but the waits of a node's slots overlap.
        int32_t got = kEmpty;
        bool resolved = false, lpoll = false;
        uint32_t pmask = 0;
        const bool wide = (iter & a.wide_mask) == 0u;
        if (active && !fresh) {
            if (need_draw) {
                const Attempt at = attempt(k0, k1, T, t, k, r, td);
                ++c_draws;
                need_draw = false;
                if (at.direct) {
                    got = node_of(at.pair);
                    resolved = true;
                } else {
                    src = at.pair;
                    have_src = true;
                }
            }
            if (BATCH && have_src && skip > 0) {
                // batches only: thousands of graphs' early rounds would
                // otherwise flood memory with re-reads
                --skip;
            } else if (have_src) {
                lpoll = true;
            }
            const bool head_ok = !(BATCH && skip > 0);
#pragma unroll
            for (int kk = 0; kk < DMAX; ++kk)
                if ((uint32_t)kk >= k && (!BATCH || (uint32_t)kk < dd) && val[kk] < kEmpty &&
                    ((uint32_t)kk == k ? head_ok : wide))
                    pmask |= 1u << kk;
        }
        const bool any_read = __ballot(lpoll || pmask != 0) != 0;
        if (any_read) {
#pragma unroll
            for (int kk = 0; kk < DMAX; ++kk) {
                const bool legacy = lpoll && (uint32_t)kk == k;
                if (legacy || ((pmask >> kk) & 1u))
                    __builtin_amdgcn_global_load_lds(
                        (const __attribute__((address_space(1))) void *)(G + (legacy ? src : pending_src(val[kk]))),
                        (__attribute__((address_space(3))) void *)(info + kk * 64), 4, 0, kSc1);
            }
            c_polls += (uint32_t)__popc(pmask) + (lpoll ? 1u : 0u);
        }
        if (fresh_mask != 0 || any_read) {
            __builtin_amdgcn_s_waitcnt(kWaitVm0);
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }

        if (fresh) {
            // attempt 0 of every slot: val holds ~node or the source pair, the
            // stage the value read from a copy source
#pragma unroll
            for (int kk = 0; kk < DMAX; ++kk) {
                if (!BATCH || (uint32_t)kk < fd) {
                    const int32_t inf = val[kk];
                    int32_t v = ~inf;
                    if (inf >= 0) {
                        v = stage[first + kk];
                        // slot values are >= -1; anything else cannot be a read
                        if (v <= kEmpty) {
                            // an empty source is remembered in the slot itself
                            ++c_pmiss;
                            v = inf != kFarPair ? pending_code((uint32_t)inf) : kEmpty;
                        }
                    }
                    val[kk] = v;
                }
            }
            c_draws += dd;
            k = 0;
            r = 0;
            skip = 0;
            misses = 0;
            need_draw = false;
            have_src = false;
            fresh = false;
        } else if (active) {
            if (lpoll) {
                const int32_t v = info[k * 64u + lane];
                if (v > kEmpty) {
                    got = v;
                    resolved = true;
                    have_src = false;
                    misses = 0;
                } else {
                    ++c_stalls;
                    // wait 0, 1, 2, 4, 8, 8, ... passes before the next re-read
                    skip = misses == 0 ? 0u : (1u << (misses < 4 ? misses - 1 : 3));
                    ++misses;
                }
            }
            if (resolved) {
                progressed = true;
#pragma unroll
                for (int kk = 0; kk < DMAX; ++kk)
                    if ((uint32_t)kk == k) val[kk] = got;
            }
#pragma unroll
            for (int kk = 0; kk < DMAX; ++kk) {
                if ((uint32_t)kk >= k && (!BATCH || (uint32_t)kk < dd) && val[kk] < kEmpty) {
                    const bool head = (uint32_t)kk == k;
                    const int32_t got_p = (pmask >> kk) & 1u ? info[kk * 64 + lane] : kEmpty;
                    if (got_p > kEmpty) {
                        val[kk] = got_p;
                        progressed = true;
                        if (head) misses = 0;
                    } else if (head) {
                        if (BATCH && skip > 0) {
                            --skip;
                        } else {
                            ++c_stalls;
                            skip = misses == 0 ? 0u : (1u << (misses < 4 ? misses - 1 : 3));
                            ++misses;
                        }
                    } else if (wide) {
                        ++c_stalls;
                    }
                }
            }
        }
        __builtin_amdgcn_wave_barrier();

        // ---- accept slots in order: duplicate check against filled slots ----
        bool finished = false, retried = false;
        if (active && !need_draw && !have_src) {
            bool stop = false;
#pragma unroll
            for (int kk = 0; kk < DMAX; ++kk) {
                if (!stop && (uint32_t)kk == k && (!BATCH || (uint32_t)kk < dd)) {
                    const int32_t c = val[kk];
                    if (c < kEmpty) {
                        // attempt 0's copy source is still empty: keep re-reading it
                        stop = true;
                    } else if (c == kEmpty) {
                        // an empty source too far out to remember: re-derive it
                        need_draw = true;
                        stop = true;
                    } else {
                        bool dup = false;
#pragma unroll
                        for (int i = 0; i < kk; ++i) dup |= (val[i] == c);
                        if (dup) {
                            ++r;
                            retried = true;
                            need_draw = true;
                            stop = true;
                            if (r >= kMaxRetry) fail = kErrRetry;
                        } else {
                            if (early) store_coherent(G + td + kk, c);
                            ++k;
                            r = 0;
                            progressed = true;
                        }
                    }
                }
            }
            if (k == dd) {
                active = false;
                finished = true;
                if (BATCH) c_slots += dd;
            }
        }
        s_retries += (uint64_t)__popcll(__ballot(retried));
        s_nodes += (uint64_t)__popcll(__ballot(finished));

        // ---- publish finished rows: coalesced through LDS (wavefront-uniform) ----
        // Each pass publishes the finished rows within 64 nodes of the smallest
        // unpublished one, as whole contiguous sectors; rows left after the
        // passes are stored by their own lanes.
        bool pending = finished && !early;
        for (int pass = 0; pass < kFlushPasses && __ballot(pending) != 0; ++pass) {
            const uint64_t key = pending ? (((uint64_t)gid << 32) | t) : ~0ull;
            const uint64_t kmin = wave_min_u64(key);
            const bool in_range = pending && key - kmin < 64;
            // bit r set <=> staged row r (node t0 + r) is finished this iteration;
            // indexed by row, not by lane
            const uint64_t rmask = wave_or_u64(in_range ? 1ull << (key - kmin) : 0ull);
            const uint32_t lead = (uint32_t)__builtin_ctzll(__ballot(pending && key == kmin));
            const uint32_t Ld = BATCH ? (uint32_t)__builtin_amdgcn_readlane((int)dd, (int)lead) : (uint32_t)DMAX;
            int32_t *const LG = BATCH ? (int32_t *)readlane64((uint64_t)G, lead) : a.G;
            const uint64_t t0 = kmin & 0xffffffffull;
            if (in_range) {
                const uint32_t row = (uint32_t)(key - kmin);
#pragma unroll
                for (int kk = 0; kk < DMAX; ++kk)
                    if (!BATCH || (uint32_t)kk < dd) stage[row * Ld + kk] = val[kk];
            }
            s_staged += (uint64_t)__popcll(rmask);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const __amdgpu_buffer_rsrc_t out = rsrc_of(LG + t0 * Ld, 64u * Ld * 4u);
            if (!BATCH && DMAX % 4 == 0) {
                // 16-byte pieces: piece p covers ints [4p, 4p + 4) of the staged rows
#pragma unroll
                for (int i = 0; i < DMAX / 4; ++i) {
                    const uint32_t p = (uint32_t)i * 64u + lane;
                    if ((rmask >> ((p * 4u) / (uint32_t)DMAX)) & 1ull) {
                        const v4u v = *reinterpret_cast<const v4u *>(&stage[p * 4u]);
                        __builtin_amdgcn_raw_buffer_store_b128(v, out, (int)(p * 16u), 0, kSc1);
                    }
                }
            } else {
                // 4-byte pieces: int e of the staged rows belongs to row e / Ld
                const uint32_t magic = (65536u + Ld - 1u) / Ld;  // exact for e < 1024, Ld <= 16
#pragma unroll
                for (int i = 0; i < DMAX; ++i) {
                    if ((uint32_t)i < Ld) {
                        const uint32_t e = (uint32_t)i * 64u + lane;
                        const uint32_t row = BATCH ? (e * magic) >> 16 : e / (uint32_t)DMAX;
                        if ((rmask >> row) & 1ull)
                            __builtin_amdgcn_raw_buffer_store_b32((uint32_t)stage[e], out, (int)(e * 4u), 0, kSc1);
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();
            pending = pending && !in_range;
        }
        if (pending) {
            if (!BATCH && DMAX % 2 == 0) {
                // rows are 8-byte aligned: 8-byte agent-scope stores
                uint64_t *row64 = reinterpret_cast<uint64_t *>(G + td);
#pragma unroll
                for (int kk = 0; kk < DMAX / 2; ++kk)
                    __hip_atomic_store(row64 + kk,
                                       ((uint64_t)(uint32_t)val[2 * kk + 1] << 32) | (uint32_t)val[2 * kk],
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else {
#pragma unroll
                for (int kk = 0; kk < DMAX; ++kk)
                    if (!BATCH || (uint32_t)kk < dd) store_coherent(G + td + kk, val[kk]);
            }
        }

        // ---- batches: per-graph completion ----
        // After this pass's rows are published, the finished nodes are added
        // to their graphs' counters, one atomic per graph present in the
        // wavefront; whoever completes a graph raises its host flag, so a
        // caller can use that graph before the whole batch ends. Every store
        // that publishes a slot is an agent-scope atomic store (sc1), which is
        // complete at agent scope once vmcnt drains, so a wait orders them
        // before the count (a release fence would also write back the L2 every
        // pass); the counter's atomic order then puts every wavefront's rows of
        // a graph before the count that completes it.
        if (BATCH && a.done_flag) {
            uint64_t fm = __ballot(finished);
            if (fm != 0) {
                __builtin_amdgcn_s_waitcnt(kWaitVm0);
                while (fm != 0) {
                    const int leader = __builtin_ctzll(fm);
                    const uint32_t g = (uint32_t)__builtin_amdgcn_readlane((int)gid, leader);
                    const uint64_t m = __ballot(finished && gid == g);
                    if ((int)lane == leader) {
                        const uint32_t c = (uint32_t)__popcll(m);
                        const uint32_t old = __hip_atomic_fetch_add(&a.done_nodes[g], c, __ATOMIC_RELAXED,
                                                                    __HIP_MEMORY_SCOPE_AGENT);
                        const BatchGraph &gr = a.graphs[g];
                        if (a.done_flag && (uint64_t)old + c == gr.n - gr.start) {
                            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                            __hip_atomic_store(&a.done_flag[g], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
                        }
                    }
                    fm &= ~m;
                }
            }
        }

        // ---- failure, watchdog and back-off (wavefront-uniform) ----
        if (__ballot(fail != 0) != 0) {
            if (fail != 0) atomicCAS(&a.scratch->status, 0, fail);
            break;
        }
        ++iter;
        if ((iter & 63u) == 0) {
            int st = 0;
            if (lane == 0) {
                st = __hip_atomic_load(&a.scratch->status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (st == 0 && wall_clock64() - start_clock > a.watchdog_ticks) {
                    atomicCAS(&a.scratch->status, 0, kErrTimeout);
                    st = kErrTimeout;
                }
            }
            if (__builtin_amdgcn_readfirstlane(st) != 0) break;
        }
        if (__ballot(progressed) == 0) {
            // Every lane of this wavefront is waiting on another node.
            ++s_sleeps;
            for (uint32_t s = 0; s < backoff; ++s) __builtin_amdgcn_s_sleep(4);
            backoff = backoff * 2 < a.sleep_cap ? backoff * 2 : a.sleep_cap;
        } else {
            backoff = 1;
        }
    }

    // ---- statistics: one atomic per counter per wavefront ----
    const uint64_t draws = wave_sum64(c_draws);
    const uint64_t slots = BATCH ? wave_sum64(c_slots) : s_nodes * (uint64_t)DMAX;
    const uint64_t pmiss = wave_sum64(c_pmiss), polls = wave_sum64(c_polls), stalls = wave_sum64(c_stalls);
    if (lane == 0) {
        unsigned long long *st = a.scratch->stats;
        const uint64_t v[11] = {s_nodes, slots, draws, s_retries, pmiss, polls, stalls, s_claims, s_sleeps,
                                s_staged, iter};
        const int idx[11] = {kStatNodes, kStatSlots,  kStatDraws,  kStatRetries, kStatPrefetchMisses, kStatPolls,
                             kStatStalls, kStatClaims, kStatSleeps, kStatStagedRows, kStatPasses};
#pragma unroll
        for (int i = 0; i < 11; ++i)
            if (v[i]) atomicAdd(&st[idx[i]], (unsigned long long)v[i]);
    }
}

// ---- out-degrees above 32 -------------------------------------------------
// A node of d > 32 slots is a whole wavefront's work: d registers per lane
// would not fit, and a node's d copy reads alone keep a wavefront's 64 lanes
// busy. The reference bounds d by its 48 KB shared-memory wait queue of 6,144
// entries per block (cuPPA_RRP_1.cu:49-56) and its int16 slot index
// (Common/DataTypes.h:22-27); here d goes up to PPA_MAX_D.
//
// A node's slots are cut into windows of 64, one slot per lane. A wavefront
// runs S streams (8, or 4 above d = 256): a stream is one window of one node.
// The wavefront claims a few consecutive nodes at a time (about one window
// per stream, ascending) and hands their windows to idle streams in order,
// so up to S windows -- of up to S nodes, or S windows of one large node --
// are in flight at once: their draws and source reads overlap, and no
// claimed node waits for the wavefront to finish the one before it. Each node
// in flight has an LDS hash set of its accepted targets (linear probing, at
// most half full), so a duplicate check costs a few LDS reads instead of a
// scan of up to d - 1 earlier slots. Only a node's earliest unfinished window
// (its "gate") accepts slots; the later ones draw and read ahead. One round:
//   1. lanes whose slot needs an attempt draw it: a direct target is known at
//      once, a copy draw reads its source slot straight into LDS (a plain read
//      first -- slots are write-once, so it returns -1 or the final value --
//      and agent-scope re-reads while the source is still empty). A slot
//      rejected before draws its next kb attempts at once (kb doubling up to
//      4 or 16) and takes the first not already in the set: a node whose t
//      candidates barely exceed its d slots (t close to d) has a coupon
//      collector's tail of thousands of attempts in its last slots, and each
//      of them would otherwise cost a memory round trip of its own;
//   2. in a gate window, the first open lane (from `acc`, the window's
//      accepted slots) whose candidate is missing, already in the set, or
//      equal to an earlier open lane's candidate ends the accepted run: the
//      lanes before it insert their candidates into the set and publish their
//      slots at once;
//   3. every open lane whose candidate is in the set (in any window), or
//      repeats an accepted one (in the gate window), draws its next attempt:
//      it is rejected whatever the slots between decide.
// The draws and the order of acceptance are the sequential model's, so the
// graph is the same. The smallest unfinished node's gate window only waits on
// finished, published slots and is always held by a stream (windows are
// handed out and finished in order), so the grid always makes progress.

__device__ __forceinline__ uint32_t set_slot(int32_t v, uint32_t hbits)
{
    return ((uint32_t)v * 0x9E3779B1u) >> (32u - hbits);  // Fibonacci hashing; hbits >= 7
}

__device__ __forceinline__ bool set_contains(const int32_t *set, uint32_t hbits, int32_t v)
{
    const uint32_t mask = (1u << hbits) - 1u;
    for (uint32_t i = set_slot(v, hbits);; i = (i + 1u) & mask) {
        const int32_t e = __hip_atomic_load(&set[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (e == v) return true;
        if (e == kEmpty) return false;  // the set is never full: the probe ends
    }
}

// v is not in the set, and the lanes inserting together hold distinct values.
__device__ __forceinline__ void set_insert(int32_t *set, uint32_t hbits, int32_t v)
{
    const uint32_t mask = (1u << hbits) - 1u;
    for (uint32_t i = set_slot(v, hbits);; i = (i + 1u) & mask) {
        int32_t none = kEmpty;
        if (__hip_atomic_compare_exchange_strong(&set[i], &none, v, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_WORKGROUP))
            return;
    }
}

// The window table of a stream: for each candidate value of a round's open
// lanes, the lowest lane holding it (128 keys, at most 64 in use, cleared per
// round).
constexpr uint32_t kWinBits = 7;
static_assert(kWideWinWords == 2u << kWinBits, "keys and lowest lanes");

// Per-lane stream flags, packed with the attempt batch size kb (bits 0..7).
constexpr uint32_t kKbMask = 0xffu;
constexpr uint32_t kFNeed = 1u << 8;    // the slot's next kb attempts are to be drawn
constexpr uint32_t kFPend = 1u << 9;    // waiting on the copy source src
constexpr uint32_t kFHave = 1u << 10;   // cand holds the current attempt's target
constexpr uint32_t kFChk = 1u << 11;    // cand is known not to be in the node's set
constexpr uint32_t kFInSet = 1u << 12;  // this round: cand is in the set
constexpr uint32_t kFInWin = 1u << 13;  // this round: cand is entered in the window table

// The slot of v in the window table, inserting it if absent.
__device__ __forceinline__ uint32_t win_insert(int32_t *keys, int32_t v)
{
    const uint32_t mask = (1u << kWinBits) - 1u;
    for (uint32_t i = set_slot(v, kWinBits);; i = (i + 1u) & mask) {
        int32_t seen = kEmpty;
        if (__hip_atomic_compare_exchange_strong(&keys[i], &seen, v, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_WORKGROUP) ||
            seen == v)
            return i;
    }
}

// One wavefront per workgroup, S streams. Dynamic LDS (int32 words): nslots
// hash sets of 1 << hbits words (at least 2d each), S window tables (keys,
// then lowest lanes), then S groups of srows rows of 64 words: attempt r + j
// of lane l of stream s lands in word l of row j of group s, and a re-read of
// its pending source in row 0.
template <int S>
__global__ __launch_bounds__(64) void wide_generate_kernel(GenArgs a, uint32_t d, uint32_t hbits, uint32_t srows,
                                                           uint32_t prows, uint32_t nslots, double inv_d)
{
    extern __shared__ int32_t wide_lds[];
    const uint32_t H = 1u << hbits;
    int32_t *const sets = wide_lds;
    int32_t *const wkey = sets + nslots * H;
    uint32_t *const wlane = reinterpret_cast<uint32_t *>(wkey + S * (1u << kWinBits));
    int32_t *const spec = wkey + S * kWideWinWords;
    int32_t *const pool = spec + S * srows * kWave;  // prows more rows for the oldest node's gate
    // row j of stream s: its own srows rows, then the pool's
    auto row = [&](int s, uint32_t j) {
        return j < srows ? spec + ((uint32_t)s * srows + j) * kWave : pool + (j - srows) * kWave;
    };
    const uint32_t kb_max = srows + prows;
    const uint32_t lane = lane_id();
    const uint32_t nwin = (d + kWave - 1) / kWave;
    const uint64_t start_clock = wall_clock64();
    for (uint32_t i = lane; i < S * (1u << kWinBits); i += kWave) {
        wkey[i] = kEmpty;
        wlane[i] = kWave;
    }
    __syncthreads();

    // Wavefront-uniform hand-out state: claimed nodes [nt, ce), window nw of
    // node nt next; cur_slot is node nt's hash set once its first window is out.
    uint64_t nt = 0, ce = 0;
    uint32_t nw = 0, cur_slot = 0, free_slots = (1u << nslots) - 1u;
    bool claims_done = false;
    // Wavefront-uniform stream state: its node (n < 2^31), and packed: in use
    // (bit 0), hash set (bits 1..3), window (bits 4..11), accepted slots of
    // the window (bits 12..18).
    uint32_t st_t[S], st_x[S];
    auto on = [&](int s) { return (st_x[s] & 1u) != 0; };
    auto st_ns = [&](int s) { return (st_x[s] >> 1) & 7u; };
    auto st_w = [&](int s) { return (st_x[s] >> 4) & 0xffu; };
    auto st_acc = [&](int s) { return (st_x[s] >> 12) & 0x7fu; };
    auto width_of = [&](int s) {
        const uint32_t rest = d - st_w(s) * kWave;
        return rest < (uint32_t)kWave ? rest : (uint32_t)kWave;
    };
    // Per-lane stream state: the slot's candidate, its current attempt (or the
    // next one to draw), flags and kb, window-table slot, pending source.
    int32_t cand[S];
    uint32_t rr[S], fl[S], wslot[S];
    uint64_t src[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        st_t[s] = st_x[s] = 0;
        cand[s] = kEmpty;
        rr[s] = fl[s] = wslot[s] = 0;
        src[s] = 0;
    }
    uint64_t s_nodes = 0, s_claims = 0, s_sleeps = 0, s_rounds = 0;
    uint64_t c_draws = 0, c_retries = 0;
    uint32_t c_pmiss = 0, c_polls = 0, c_stalls = 0;
    uint32_t backoff = 1;
    int fail = 0;

    while (true) {
        bool progressed = false;
        // ---- 1. idle streams take the next windows, in order ----
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if (on(s)) continue;
            if (nw == 0 && nt >= ce && !claims_done) {
                uint64_t base = 0;
                if (lane == 0) base = atomicAdd(&a.scratch->claim, (unsigned long long)a.chunk);
                base = uniform64(base);
                ++s_claims;
                if (base >= a.n) {
                    claims_done = true;
                } else {
                    nt = base;
                    ce = base + a.chunk < a.n ? base + a.chunk : a.n;
                }
            }
            // a new node needs a free hash set; the windows of node nt follow in order
            if (nt < ce && (nw != 0 || free_slots != 0)) {
                if (nw == 0) {
                    cur_slot = (uint32_t)__builtin_ctz(free_slots);
                    free_slots &= free_slots - 1u;
                    int32_t *const set = sets + cur_slot * H;
                    for (uint32_t i = lane; i < H; i += kWave) set[i] = kEmpty;
                }
                st_t[s] = (uint32_t)nt;
                st_x[s] = 1u | cur_slot << 1 | nw << 4;
                const uint32_t width = width_of(s);
                cand[s] = kEmpty;
                rr[s] = 0;
                fl[s] = lane < width ? (kFNeed | 1u) : 1u;
                progressed = true;
                if (++nw == nwin) {
                    nw = 0;
                    ++nt;
                }
            }
        }
        bool any = false;
#pragma unroll
        for (int s = 0; s < S; ++s) any |= on(s);
        if (!any && claims_done) break;

        // ---- gates: a stream is its node's gate when no stream holds an
        // earlier window of the same node (earlier windows were handed out
        // first, so they are finished); the gate of the oldest node may use
        // the pool rows ----
        uint32_t gates = 0, pool_owner = S, pool_t = 0xffffffffu;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            bool g = on(s);
#pragma unroll
            for (int q = 0; q < S; ++q)
                if (q != s && on(q) && st_t[q] == st_t[s] && st_w(q) < st_w(s)) g = false;
            gates |= g ? 1u << s : 0u;
            if (g && st_t[s] < pool_t) {
                pool_owner = s;
                pool_t = st_t[s];
            }
        }

        // ---- 2. draws (attempt r + j into row j) and re-reads (row 0), all in flight at once ----
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if (!on(s)) continue;
            const uint64_t t = st_t[s], td = t * (uint64_t)d;
            const uint32_t k = st_w(s) * kWave + lane;
            const uint32_t cap = (uint32_t)s == pool_owner ? kb_max : srows;
            uint32_t kb = fl[s] & kKbMask;
            kb = kb < cap ? kb : cap;
            fl[s] = (fl[s] & ~kKbMask) | kb;  // what is drawn now is what step 3 reads
            const bool need = (fl[s] & kFNeed) != 0;
            for (uint32_t j = 0; j < cap; ++j) {
                const bool dj = need && j < kb;
                if (__ballot(dj) == 0) break;
                if (dj) {
                    const Attempt at = attempt(a.key0, a.key1, a.T, t, k, rr[s] + j, td);
                    ++c_draws;
                    int32_t *const rj = row(s, j);
                    if (at.direct)
                        rj[lane] = (int32_t)div_small(at.pair, d, inv_d);
                    else
                        __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)(a.G + at.pair),
                                                         (__attribute__((address_space(3))) void *)rj, 4, 0, 0);
                }
            }
            if (fl[s] & kFPend) {
                __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)(a.G + src[s]),
                                                 (__attribute__((address_space(3))) void *)row(s, 0), 4, 0, kSc1);
                ++c_polls;
            }
        }
        __builtin_amdgcn_s_waitcnt(kWaitVm0);
        __syncthreads();

        // ---- 3. the first drawn attempt not already rejected by the set; resolved re-reads ----
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if (!on(s)) continue;
            const int32_t *const set = sets + st_ns(s) * H;
            uint32_t f = fl[s];
            if (f & kFNeed) {
                f &= ~kFNeed;
                progressed = true;
                const uint32_t kb = f & kKbMask;
                uint32_t j = 0;
                bool done = false;
                while (!done && j < kb) {
                    if (rr[s] + j >= kMaxRetry) {
                        fail = kErrRetry;
                        done = true;
                        break;
                    }
                    const int32_t v = row(s, j)[lane];
                    if (v <= kEmpty) {
                        // its copy source is still empty: re-derive the source and wait on it
                        ++c_pmiss;
                        src[s] = attempt(a.key0, a.key1, a.T, st_t[s], st_w(s) * kWave + lane, rr[s] + j,
                                         (uint64_t)st_t[s] * d)
                                     .pair;
                        f |= kFPend;
                        done = true;
                    } else if (set_contains(set, hbits, v)) {
                        ++c_retries;  // it repeats an accepted slot: rejected
                        ++j;
                    } else {
                        cand[s] = v;
                        f |= kFHave | kFChk;
                        done = true;
                    }
                }
                rr[s] += j;
                if (!done) f = (f & ~kKbMask) | kFNeed | (kb * 2 < kb_max ? kb * 2 : kb_max);
            } else if (f & kFPend) {
                const int32_t v = row(s, 0)[lane];
                if (v > kEmpty) {
                    cand[s] = v;
                    f = (f & ~(kFPend | kFChk)) | kFHave;
                    progressed = true;
                } else {
                    ++c_stalls;
                }
            }
            fl[s] = f;
        }

        // ---- 4. set lookups, window tables (gates as found above) ----
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if (!on(s)) continue;
            const bool gate = (gates >> s) & 1u;
            const int32_t *const set = sets + st_ns(s) * H;
            const bool open = lane < width_of(s) && lane >= st_acc(s);
            uint32_t f = fl[s] & ~(kFInSet | kFInWin);
            const bool have = (f & kFHave) != 0;
            // a window behind the gate looks every candidate up again: the
            // gate may have added to the set since
            const bool chk = gate && (f & kFChk) != 0;
            if (open && have && !chk && set_contains(set, hbits, cand[s])) f |= kFInSet;
            if (gate && open && have && !(f & kFInSet)) {
                f |= kFInWin;
                wslot[s] = win_insert(wkey + (uint32_t)s * (1u << kWinBits), cand[s]);
                __hip_atomic_fetch_min(&wlane[(uint32_t)s * (1u << kWinBits) + wslot[s]], lane, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
            }
            fl[s] = f;
        }
        __syncthreads();

        // ---- 5. accept in order in the gates; reject what is decided ----
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if (!on(s)) continue;
            const bool gate = (gates >> s) & 1u;
            int32_t *const set = sets + st_ns(s) * H;
            const uint32_t width = width_of(s);
            const uint32_t acc = st_acc(s);
            const bool open = lane < width && lane >= acc;
            uint32_t f = fl[s];
            const bool have = (f & kFHave) != 0, in_set = (f & kFInSet) != 0, in_win = (f & kFInWin) != 0;
            bool reject;
            uint32_t first = acc;
            if (gate) {
                const uint32_t ws = (uint32_t)s * (1u << kWinBits) + wslot[s];
                const uint32_t owner =
                    in_win ? __hip_atomic_load(&wlane[ws], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) : lane;
                // the accepted run cannot pass the first open lane still without a candidate
                const uint64_t wait_mask = __ballot(open && !have);
                const uint32_t lim = wait_mask ? (uint32_t)__builtin_ctzll(wait_mask) : width;
                const uint64_t bad_mask = __ballot(open && lane < lim && (in_set || owner < lane));
                first = bad_mask ? (uint32_t)__builtin_ctzll(bad_mask) : lim;
                if (open && lane < first) {
                    set_insert(set, hbits, cand[s]);
                    store_coherent(a.G + (uint64_t)st_t[s] * d + st_w(s) * kWave + lane, cand[s]);
                }
                if (in_win) {
                    // every lane has read its owner above: clear the table for the next round
                    __hip_atomic_store(&wkey[ws], kEmpty, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    __hip_atomic_store(&wlane[ws], (uint32_t)kWave, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                }
                if (first > acc) progressed = true;
                // Every open slot from `first` on whose candidate is in the set
                // or repeats an accepted slot's is rejected whatever the slots
                // between decide; the others stay checked: their candidates are
                // neither in the set nor among the values just added to it.
                reject = open && lane >= first && have && (in_set || owner < first);
                st_x[s] = (st_x[s] & ~(0x7fu << 12)) | first << 12;
            } else {
                reject = open && have && in_set;
            }
            if (reject) {
                const uint32_t kb = f & kKbMask;
                f = (f & ~(kFHave | kFChk | kKbMask)) | kFNeed | (kb * 2 < kb_max ? kb * 2 : kb_max);
                ++c_retries;
                progressed = true;
                if (++rr[s] >= kMaxRetry) fail = kErrRetry;
            }
            f = gate && (f & kFHave) ? (f | kFChk) : (f & ~kFChk);
            fl[s] = f;
            if (gate && first == width) {
                // window done: the stream is free; the node is done with its last window
                if (st_w(s) == nwin - 1u) {
                    ++s_nodes;
                    free_slots |= 1u << st_ns(s);
                }
                st_x[s] = 0;
            }
        }

        ++s_rounds;
        if (__ballot(fail != 0) != 0) {
            if (fail != 0) atomicCAS(&a.scratch->status, 0, fail);
            break;
        }
        if ((s_rounds & 63u) == 0) {
            int st = 0;
            if (lane == 0) {
                st = __hip_atomic_load(&a.scratch->status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (st == 0 && wall_clock64() - start_clock > a.watchdog_ticks) {
                    atomicCAS(&a.scratch->status, 0, kErrTimeout);
                    st = kErrTimeout;
                }
            }
            if (__builtin_amdgcn_readfirstlane(st) != 0) break;
        }
        if (__ballot(progressed) == 0) {
            // every open slot of every stream waits on another node's slot
            ++s_sleeps;
            for (uint32_t i = 0; i < backoff; ++i) __builtin_amdgcn_s_sleep(4);
            backoff = backoff * 2 < a.sleep_cap ? backoff * 2 : a.sleep_cap;
        } else {
            backoff = 1;
        }
    }

    const uint64_t draws = wave_sum64(c_draws), retries = wave_sum64(c_retries);
    const uint64_t pmiss = wave_sum64(c_pmiss), polls = wave_sum64(c_polls), stalls = wave_sum64(c_stalls);
    if (lane == 0) {
        unsigned long long *st = a.scratch->stats;
        const uint64_t v[10] = {s_nodes, s_nodes * d, draws, retries, pmiss, polls, stalls, s_claims, s_sleeps,
                                s_rounds};
        const int idx[10] = {kStatNodes,  kStatSlots,  kStatDraws,  kStatRetries, kStatPrefetchMisses,
                             kStatPolls,  kStatStalls, kStatClaims, kStatSleeps,  kStatPasses};
#pragma unroll
        for (int i = 0; i < 10; ++i)
            if (v[i]) atomicAdd(&st[idx[i]], (unsigned long long)v[i]);
    }
}

// ---- LDS-resident prefix ----------------------------------------------------
// A small graph's run time is its longest dependency chain times the cost of
// one dependent step, and most of that chain lies among the first nodes,
// whose copy sources are nearly all still being generated. One workgroup
// generates nodes d+1 .. prefix-1 with the graph held in LDS, where a
// dependent step costs an LDS read instead of an HBM round trip, then writes
// the rows out; the persistent generator starts its claims at `prefix`.
// Same draws, same order of acceptance, hence the same graph.
//
// The workgroup's lanes form groups of GS >= d lanes (GS a power of two: a
// group of up to 16 lanes lies inside one 16-lane DPP row, one of 32 inside
// half a wavefront); a group owns nodes d+1+g,
// d+1+g+groups, ... in ascending order, one slot per lane. Attempt 0 of every
// slot (and, in a 64 KB prefix, attempt 1 for a first duplicate) is drawn
// before the first round, in parallel, and parked in the empty slot; on
// taking a node each slot lane picks it up there. A round of the group
// compares every slot's candidate with those of the slots before it (DPP row
// shifts) and finds the first slot from `acc` (the slots accepted so far) on
// that is still empty or a duplicate: the slots before it are accepted at
// once; a duplicate there, and every later slot that repeats an accepted one,
// draws its next attempt and the group goes on; an empty source ends the
// group's rounds. A slot's work is thus spread over the node's lanes instead
// of a loop over the node's slots on one lane. There are no barriers: a
// wavefront re-reads its empty sources until one fills, and LDS makes a
// stored slot visible to the others at once. The smallest unfinished node
// only depends on finished slots, so its group always progresses.
constexpr int kPrefixThreads = (int)kPrefixLanes;
constexpr int32_t kNoCandidate = -3;  // DPP fill for lanes outside the row: equals no candidate

// A parked draw in B+1 bits: bit B marks a direct link to the node in bits
// 0..B-1, else those bits are the source pair (both below the prefix's
// slots: a 64 KB prefix parks two attempts of 15 bits, a 160 KB one one of 17).
template <uint32_t B>
__device__ __forceinline__ uint32_t park(const Attempt &at, uint32_t d)
{
    // pair < P*d <= the prefix's slots: a 32-bit quotient
    return at.direct ? (1u << B) | ((uint32_t)at.pair / d) : (uint32_t)at.pair;
}

// The kernel's register form: ~node (direct, < 0) or the source pair (>= 0).
template <uint32_t B>
__device__ __forceinline__ int32_t unpark(uint32_t c)
{
    return c & (1u << B) ? ~(int32_t)(c & ((1u << B) - 1u)) : (int32_t)c;
}

// The smallest lane of this lane's group, below lane j, that holds v (j if
// none). In groups of up to 16 lanes row_shr:s hands lane j the value of lane
// j-s of its 16-lane row (kNoCandidate past the row's start); 32-lane groups
// span two rows and read lane j-s through ds_bpermute instead (lanes j < s
// read another group's lane and are masked). The GS-1 shifts are
// independent and their results meet in a tree of minima, keeping the
// dependent chain short.
template <int GS, int S>
__device__ __forceinline__ uint32_t same_at(int32_t v, uint32_t j)
{
    int32_t o;
    if constexpr (GS <= 16)
        o = __builtin_amdgcn_update_dpp(kNoCandidate, v, 0x110 + S, 0xf, 0xf, false);
    else
        o = __builtin_amdgcn_ds_bpermute((int)((lane_id() - (uint32_t)S) << 2), v);
    return j >= (uint32_t)S && o == v ? j - (uint32_t)S : j;
}

template <int GS, int... S>
__device__ __forceinline__ uint32_t first_same_of(int32_t v, uint32_t j, std::integer_sequence<int, S...>)
{
    constexpr int N = (int)sizeof...(S) + 1;
    uint32_t m[N] = {j, same_at<GS, S + 1>(v, j)...};
#pragma unroll
    for (int w = 1; w < N; w *= 2)
#pragma unroll
        for (int i = 0; i + w < N; i += 2 * w) m[i] = m[i] < m[i + w] ? m[i] : m[i + w];
    return m[0];
}

template <int GS>
__device__ __forceinline__ uint32_t first_same(int32_t v, uint32_t j)
{
    return first_same_of<GS>(v, j, std::make_integer_sequence<int, GS - 1>{});
}

template <int GS, bool BATCH, uint32_t SLOTS>
__global__ __launch_bounds__(kPrefixThreads) void prefix_kernel(PrefixArgs a)
{
    static_assert(GS >= 1 && GS <= 32 && (GS & (GS - 1)) == 0, "a group lies inside half a wavefront");
    // two parked attempts of 15 bits, or one of 17
    constexpr bool kTwo = SLOTS <= (1u << 14);
    constexpr uint32_t kB = kTwo ? 14u : 16u;
    static_assert(SLOTS <= (1u << 16), "parked draws must fit 16 bits");
    __shared__ int32_t Gs[SLOTS];
    __shared__ int32_t s_abort;  // 0, or the error that ends the prefix
    // a batch's graph, or the single graph of the arguments
    uint32_t gi = 0;
    if (BATCH) {
        gi = a.pidx[blockIdx.x];
        const BatchGraph &gr = a.graphs[gi];
        a.G = gr.G;
        a.T = gr.T;
        a.key0 = gr.key0;
        a.key1 = gr.key1;
        a.model = gr.model;
        a.d = gr.d;
        a.prefix = gr.start;
    }
    const uint32_t d = a.d;
    const uint64_t P = a.prefix;
    if (P * d > SLOTS) {
        // a prefix sized for another instance: refuse rather than overrun LDS
        if (threadIdx.x == 0) atomicCAS(&a.scratch->status, 0, kErrInternal);
        return;
    }
    const uint32_t seed_slots = (d + 1) * d;
    const uint32_t used = (uint32_t)P * d;
    // Attempt 0 of every generated slot (and attempt 1, if kTwo) is drawn up
    // front, by all lanes at once, and parked in the slot itself as bit 31 |
    // attempt 1 << 15 | attempt 0 (or bit 31 | attempt 0): negative, so
    // readers still find the slot empty.
    uint32_t draws = 0;
    for (uint32_t s = threadIdx.x; s < used; s += kPrefixThreads) {
        int32_t v;
        if (s < seed_slots) {
            v = seed_target(a.model, d, s / d, s % d);
        } else {
            const uint32_t tn = s / d, k = s - tn * d;
            const Attempt a0 = attempt(a.key0, a.key1, a.T, tn, k, 0u, (uint64_t)tn * d);
            if constexpr (kTwo) {
                const Attempt a1 = attempt(a.key0, a.key1, a.T, tn, k, 1u, (uint64_t)tn * d);
                v = (int32_t)(0x80000000u | park<kB>(a1, d) << 15 | park<kB>(a0, d));
            } else {
                v = (int32_t)(0x80000000u | park<kB>(a0, d));
            }
            ++draws;
        }
        Gs[s] = v;
    }
    if (threadIdx.x == 0) s_abort = 0;
    __syncthreads();

    constexpr uint32_t kGroups = (uint32_t)kPrefixThreads / GS;
    constexpr uint32_t kGroupMask = GS == 32 ? 0xffffffffu : (1u << GS) - 1u;
    const uint32_t j = threadIdx.x % GS;                 // this lane's slot
    const uint32_t gshift = (threadIdx.x & 63u) - j;     // the group's first lane in the wavefront
    const bool slot_lane = j < d;
    // Group state, the same on all its lanes: node t, its accepted slots acc.
    uint64_t t = (uint64_t)d + 1 + threadIdx.x / GS;
    bool active = t < P;
    uint32_t acc = 0;
    // This slot's attempt r: code = ~node (direct) or its source pair (< P*d,
    // so >= 0); cv = its candidate value once known, negative while the
    // source is still empty (kEmpty, or the source's own parked draws).
    uint32_t r = 0;
    int32_t code = 0, cv = kEmpty;
    auto lds_slot = [&](int32_t c) -> int32_t {
        return c < 0 ? ~c : __hip_atomic_load(&Gs[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    };
    uint32_t nodes = 0, retries = 0;
    int32_t spare = 0;  // attempt 1, parked with attempt 0 (kTwo)
    auto take = [&]() {  // this lane's slot of node t: attempt 0, from where it is parked
        const uint32_t e =
            (uint32_t)__hip_atomic_load(&Gs[t * d + j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if constexpr (kTwo) {
            code = unpark<kB>(e & 0x7fffu);
            spare = unpark<kB>(e >> 15 & 0x7fffu);
        } else {
            code = unpark<kB>(e & 0x1ffffu);
        }
        cv = lds_slot(code);
    };
    auto draw = [&]() {  // attempt r >= 1
        if (kTwo && r == 1) {
            code = spare;
        } else {
            const Attempt at = attempt(a.key0, a.key1, a.T, t, j, r, t * d);
            code = unpark<kB>(park<kB>(at, d));
        }
        cv = lds_slot(code);
        ++draws;
    };
    if (active && slot_lane) take();
    const uint64_t start_clock = wall_clock64();
    for (uint32_t iter = 1;; ++iter) {
        // re-read the empty sources; the rounds run once one of them is filled
        bool got = false;
        if (active && slot_lane && j >= acc && cv < 0 && r < kMaxRetry) {
            cv = lds_slot(code);
            got = cv >= 0;
        }
        if (iter == 1 || __ballot(got) != 0) {
            // rounds, until no group of the wavefront can go on
            for (;;) {
                const uint32_t same = first_same<GS>(cv, j);
                const bool open = active && slot_lane && j >= acc;
                const uint64_t pm = __ballot(open && cv < 0);
                const uint64_t dm = __ballot(open && cv >= 0 && same < j);
                const uint32_t gm = (uint32_t)((pm | dm) >> gshift) & kGroupMask;
                const uint32_t first = gm ? (uint32_t)__builtin_ctz(gm) : d;
                if (open && j < first)
                    __hip_atomic_store(&Gs[t * d + j], cv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                bool more = false, fresh = false, redraw = false;
                if (active) {
                    if (first >= d) {
                        // node done: take the group's next one
                        if (j == 0) ++nodes;
                        t += kGroups;
                        active = t < P;
                        acc = 0;
                        r = 0;
                        fresh = active && slot_lane;
                        more = active;
                    } else {
                        acc = first;
                        // the group goes on after a duplicate, not after an empty source
                        more = ((pm >> (gshift + first)) & 1u) == 0;
                        // every slot from `first` on whose candidate repeats an
                        // accepted slot is rejected whatever comes between:
                        // all of them draw again now
                        if (open && cv >= 0 && j >= first && same < first) {
                            ++retries;
                            if (++r >= kMaxRetry) {
                                cv = kEmpty;  // stays empty: the prefix is abandoned below
                            } else {
                                redraw = true;
                            }
                        }
                    }
                }
                if (fresh) take();
                if (redraw) draw();
                if (__ballot(more) == 0) break;
            }
        } else {
            __builtin_amdgcn_s_sleep(1);
        }
        if (__ballot(active) == 0) break;
        // A lane past the retry cap (whose slot others may be waiting for) or
        // the wall-clock budget ends the prefix for every wavefront.
        const bool capped = __ballot(r >= kMaxRetry) != 0;
        if (lane_id() == 0) {
            int32_t err = capped ? kErrRetry : 0;
            if (err == 0 && (iter & 63u) == 0 && wall_clock64() - start_clock > a.watchdog_ticks) err = kErrTimeout;
            if (err != 0) {
                int32_t none = 0;
                __hip_atomic_compare_exchange_strong(&s_abort, &none, err, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
        if (__hip_atomic_load(&s_abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) != 0) break;
    }
    __syncthreads();
    const int32_t fail = s_abort;
    const bool failed = fail != 0;
    // (a failed prefix leaves parked draws: written out as empty slots)
    for (uint32_t s = seed_slots + threadIdx.x; s < used; s += kPrefixThreads) {
        const int32_t v = Gs[s];
        a.G[s] = v < 0 ? kEmpty : v;
    }
    if (failed && threadIdx.x == 0) atomicCAS(&a.scratch->status, 0, fail);
    if (BATCH && a.done_flag && P == a.graphs[gi].n && !failed) {
        // the prefix is the whole graph: no generator pass will raise its flag
        __syncthreads();
        if (threadIdx.x == 0) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
            __hip_atomic_store(&a.done_flag[gi], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }

    const uint32_t wnodes = wave_sum(nodes), wdraws = wave_sum(draws), wretries = wave_sum(retries);
    if (lane_id() == 0) {
        unsigned long long *st = a.scratch->stats;
        if (wnodes) {
            atomicAdd(&st[kStatNodes], (unsigned long long)wnodes);
            atomicAdd(&st[kStatSlots], (unsigned long long)wnodes * d);
        }
        if (wdraws) atomicAdd(&st[kStatDraws], (unsigned long long)wdraws);
        if (wretries) atomicAdd(&st[kStatRetries], (unsigned long long)wretries);
    }
}

// ---- degree pass ----------------------------------------------------------

__global__ void degree_init_kernel(uint32_t *deg, uint64_t n, uint32_t d)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) deg[v] = d;
}

// deg[target]++ for every slot of nodes t >= d+1 (cuPPA.cu:115-131), one
// atomic per slot: the fallback when the partitioned pass (ppa_degree.hip)
// cannot get its buffers. Out-of-range values (an incomplete graph) are
// skipped, never dereferenced.
__global__ void degree_count_kernel(const int32_t *G, uint64_t first_slot, uint64_t slots, uint64_t n,
                                    uint32_t *deg, int32_t *status)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t s = first_slot + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < slots; s += stride) {
        const int32_t v = G[s];
        if (v >= 0 && (uint64_t)v < n) check_u32_add(atomicAdd(&deg[v], 1u), 1u, status);
    }
}

// deg[v] += base_degree(model, d, v) - d for the seed nodes v < d: the
// Barabasi-Albert count (PAHelper.hpp:84-121) leaves out their self-loops.
__global__ void seed_degree_kernel(uint32_t *deg, uint32_t d, uint32_t model)
{
    for (uint32_t v = threadIdx.x; v < d; v += blockDim.x) deg[v] = deg[v] - d + base_degree(model, d, v);
}

constexpr int kHistLds = 4096;

// hist[deg[v]]++ (cuPPA.cu:147-158); small degrees are counted in LDS first.
__global__ __launch_bounds__(256) void histogram_kernel(const uint32_t *deg, uint64_t n, uint64_t *hist,
                                                        uint64_t hist_len, DeviceScratch *s)
{
    __shared__ uint32_t local[kHistLds];
    const uint32_t lds_len = hist_len < kHistLds ? (uint32_t)hist_len : kHistLds;
    for (uint32_t i = threadIdx.x; i < kHistLds; i += blockDim.x) local[i] = 0;
    __syncthreads();
    uint32_t mx = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        const uint32_t x = deg[v];
        mx = x > mx ? x : mx;
        if (hist) {
            // the most common degrees repeat across a wavefront: the lanes
            // holding the first active lane's value add as one (1.84 -> 1.57
            // ms over 2^30 degrees at d = 8; eight lane-strided LDS copies of
            // the counters instead took 2.17)
            const uint32_t x0 = __builtin_amdgcn_readfirstlane(x);
            const uint64_t same = __ballot(x == x0);
            if (x == x0) {
                if (rank_below(same) == 0) {
                    const uint32_t c = (uint32_t)__popcll(same);
                    if (x < lds_len) atomicAdd(&local[x], c);
                    else if (x < hist_len) atomicAdd((unsigned long long *)&hist[x], (unsigned long long)c);
                }
            } else if (x < lds_len) {
                atomicAdd(&local[x], 1u);
            } else if (x < hist_len) {
                atomicAdd((unsigned long long *)&hist[x], 1ull);
            }
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t o = __shfl_xor(mx, off, kWave);
        mx = o > mx ? o : mx;
    }
    if (lane_id() == 0) atomicMax(&s->max_degree, (unsigned long long)mx);
    __syncthreads();
    if (hist)
        for (uint32_t i = threadIdx.x; i < lds_len; i += blockDim.x)
            if (local[i]) atomicAdd((unsigned long long *)&hist[i], (unsigned long long)local[i]);
}

__global__ void zero_u64_kernel(uint64_t *p, uint64_t len)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += stride) p[i] = 0;
}

uint32_t grid_for(uint64_t work, uint32_t block, uint32_t cap)
{
    uint64_t g = (work + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (uint32_t)g;
}

template <int D, bool BATCH>
hipError_t launch_generate_d(const GenArgs &a, uint32_t blocks, uint32_t threads, hipStream_t stream)
{
    hipLaunchKernelGGL((generate_kernel<D, BATCH>), dim3(blocks), dim3(threads), 0, stream, a);
    return hipGetLastError();
}

template <int GS>
hipError_t launch_prefix_gs(PrefixArgs a, uint32_t d, hipStream_t stream)
{
    // the host sized the prefix by prefix_slots(d, true): the 160 KB kernel for 16- and 32-lane groups
    static_assert(GS >= 16 || GS < (int)kWidePrefixMinD, "16- and 32-lane groups are exactly d >= kWidePrefixMinD");
    if (a.prefix * d > (GS >= 16 ? kPrefixSlotsWide : kPrefixSlots)) return hipErrorInvalidValue;
    a.d = d;
    if constexpr (GS >= 16)
        hipLaunchKernelGGL((prefix_kernel<GS, false, kPrefixSlotsWide>), dim3(1), dim3(kPrefixThreads), 0, stream, a);
    else
        hipLaunchKernelGGL((prefix_kernel<GS, false, kPrefixSlots>), dim3(1), dim3(kPrefixThreads), 0, stream, a);
    return hipGetLastError();
}

template <int D, bool BATCH>
hipError_t occupancy_d(uint32_t threads, int *out)
{
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(out, generate_kernel<D, BATCH>, (int)threads, 0);
}

hipError_t launch_generate_wide(uint32_t d, const GenArgs &a, uint32_t blocks, uint32_t threads,
                                hipStream_t stream)
{
    // a.chunk nodes per claim: zero would never advance the claims
    if (d <= kNarrowMaxD || d > PPA_MAX_D || threads != kWideThreads || a.chunk == 0) return hipErrorInvalidValue;
    const WideGeometry g = wide_geometry(d);
    // kb is kept in 8 bits; a stream holds at most 64 lanes' windows of 64 slots
    if (g.srows + g.prows > 255 || g.nslots > 8 || g.nslots < 1) return hipErrorInvalidValue;
#define PPA_WIDE(SS)                                                                                            \
    case SS:                                                                                                    \
        hipLaunchKernelGGL(wide_generate_kernel<SS>, dim3(blocks), dim3(kWideThreads), wide_lds_bytes(g), stream, \
                           a, d, g.hbits, g.srows, g.prows, g.nslots, 1.0 / d);                                 \
        break;
    switch (g.streams) {
        PPA_WIDE(1) PPA_WIDE(2) PPA_WIDE(4) PPA_WIDE(8)
    default: return hipErrorInvalidValue;
    }
#undef PPA_WIDE
    return hipGetLastError();
}

hipError_t wide_occupancy(uint32_t d, int *out)
{
    const WideGeometry g = wide_geometry(d);
    switch (g.streams) {
#define PPA_WIDE(SS) \
    case SS: return hipOccupancyMaxActiveBlocksPerMultiprocessor(out, wide_generate_kernel<SS>, (int)kWideThreads, \
                                                                 wide_lds_bytes(g));
        PPA_WIDE(1) PPA_WIDE(2) PPA_WIDE(4) PPA_WIDE(8)
#undef PPA_WIDE
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t launch_reset_scratch(DeviceScratch *s, uint64_t first, hipStream_t stream)
{
    hipLaunchKernelGGL(reset_scratch_kernel, dim3(1), dim3(64), 0, stream, s, first);
    return hipGetLastError();
}

hipError_t launch_fill(const FillArgs &a, hipStream_t stream)
{
    // one block per span: the 16-byte groups (and the partial last one), then
    // the n + 1 row offsets
    const uint64_t slot_blocks = ((a.n * a.d) / 4 + 1 + kFillSpan - 1) / kFillSpan;
    const uint64_t row_blocks = a.row_ptr ? (a.n + 1 + kFillSpan - 1) / kFillSpan : 0;
    const uint64_t blocks = slot_blocks + row_blocks;
    const uint64_t grid = blocks < (1ull << 30) ? blocks : (1ull << 30);
    hipLaunchKernelGGL(fill_kernel, dim3((uint32_t)grid), dim3(kFillThreads), 0, stream, a, slot_blocks, blocks);
    return hipGetLastError();
}

hipError_t launch_fill_batch(const BatchFillArgs &a, uint64_t total_slots, uint64_t total_rows,
                             hipStream_t stream)
{
    const uint64_t work = (total_slots > total_rows ? total_slots : total_rows) / kFillPerThread + 1;
    hipLaunchKernelGGL(fill_batch_kernel, dim3(grid_for(work, 256, 16384)), dim3(256), 0, stream, a,
                       total_slots, total_rows);
    return hipGetLastError();
}

#define PPA_FOR_EACH_D(X) \
    X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) \
    X(17) X(18) X(19) X(20) X(21) X(22) X(23) X(24) X(25) X(26) X(27) X(28) X(29) X(30) X(31) X(32)

hipError_t launch_generate(uint32_t d, const GenArgs &a, uint32_t blocks, uint32_t threads,
                           hipStream_t stream)
{
    switch (d) {
#define PPA_CASE(DD) \
    case DD: return launch_generate_d<DD, false>(a, blocks, threads, stream);
        PPA_FOR_EACH_D(PPA_CASE)
#undef PPA_CASE
    default: return launch_generate_wide(d, a, blocks, threads, stream);
    }
}

hipError_t launch_prefix(uint32_t d, const PrefixArgs &a, hipStream_t stream)
{
    // groups of the next power of two >= d lanes
    if (d == 0 || d > kPrefixMaxD) return hipErrorInvalidValue;
    if (d == 1) return launch_prefix_gs<1>(a, d, stream);
    if (d == 2) return launch_prefix_gs<2>(a, d, stream);
    if (d <= 4) return launch_prefix_gs<4>(a, d, stream);
    if (d <= 8) return launch_prefix_gs<8>(a, d, stream);
    if (d <= 16) return launch_prefix_gs<16>(a, d, stream);
    return launch_prefix_gs<32>(a, d, stream);
}

hipError_t launch_prefix_batch(const PrefixArgs &a, uint32_t count, hipStream_t stream)
{
    static_assert(kBatchDMax == 16, "batch prefix groups are 16 lanes");
    hipLaunchKernelGGL((prefix_kernel<16, true, kPrefixSlots>), dim3(count), dim3(kPrefixThreads), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_generate_batch(const GenArgs &a, uint32_t blocks, uint32_t threads, hipStream_t stream)
{
    return launch_generate_d<kBatchDMax, true>(a, blocks, threads, stream);
}

hipError_t generate_occupancy(uint32_t d, uint32_t threads, int *out)
{
    switch (d) {
    case 0: return occupancy_d<kBatchDMax, true>(threads, out);
#define PPA_CASE(DD) \
    case DD: return occupancy_d<DD, false>(threads, out);
        PPA_FOR_EACH_D(PPA_CASE)
#undef PPA_CASE
    default:
        if (d > PPA_MAX_D || threads != kWideThreads) return hipErrorInvalidValue;
        return wide_occupancy(d, out);
    }
}

hipError_t launch_degree_atomic(const int32_t *G, uint64_t n, uint32_t d, uint32_t *deg, int32_t *status,
                                hipStream_t stream)
{
    hipLaunchKernelGGL(degree_init_kernel, dim3(grid_for(n, 256, 8192)), dim3(256), 0, stream, deg, n, d);
    const uint64_t first = (uint64_t)(d + 1) * d, slots = n * d;
    if (slots > first)
        hipLaunchKernelGGL(degree_count_kernel, dim3(grid_for(slots - first, 256, 16384)), dim3(256), 0,
                           stream, G, first, slots, n, deg, status);
    return hipGetLastError();
}

hipError_t launch_seed_degree(uint32_t *deg, uint32_t d, uint32_t model, hipStream_t stream)
{
    if (model != PPA_MODEL_COPY)
        hipLaunchKernelGGL(seed_degree_kernel, dim3(1), dim3(256), 0, stream, deg, d, model);
    return hipGetLastError();
}

hipError_t launch_degree_histogram(const uint32_t *deg, uint64_t n, uint64_t *hist, uint64_t hist_len,
                                   DeviceScratch *s, hipStream_t stream)
{
    if (hist && hist_len)
        hipLaunchKernelGGL(zero_u64_kernel, dim3(grid_for(hist_len, 256, 4096)), dim3(256), 0, stream, hist,
                           hist_len);
    hipLaunchKernelGGL(histogram_kernel, dim3(grid_for(n, 256, 2048)), dim3(256), 0, stream, deg, n, hist,
                       hist_len, s);
    return hipGetLastError();
}

}  // namespace ppa
