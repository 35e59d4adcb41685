// ppa_kernels.hpp -- launch interface between the C ABI and the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ppa {

enum StatIndex : int {
    kStatNodes = 0,
    kStatSlots,
    kStatDraws,
    kStatRetries,
    kStatPrefetchMisses,
    kStatPolls,
    kStatStalls,
    kStatClaims,
    kStatSleeps,
    kStatStagedRows,
    kStatPasses,
    kStatCount
};

// Device scratch of one context (one call in flight per context).
struct DeviceScratch {
    unsigned long long claim;  // next unclaimed node (single graph) or work index (batch)
    int status;                // 0 = running/ok, else a negative PPA_ERR_* code
    int pad;
    unsigned long long stats[kStatCount];
    unsigned long long max_degree;
};

// One graph of a batch, as the kernels read it (64 bytes).
struct BatchGraph {
    int32_t *G;               // this graph's targets, row-major [n][d]
    int64_t *row_ptr;         // this graph's row offsets [n+1], nullable
    uint64_t n;
    uint64_t T;               // coin threshold
    uint64_t targets_offset;  // added to every row_ptr entry
    uint32_t key0, key1;      // Philox key
    uint32_t d;
    uint32_t model;           // PPA_MODEL_* (seed rows)
    double inv_d;             // 1.0 / d
    uint64_t start;           // first node the persistent generator claims (after the LDS prefix)
    uint64_t pad;
};
static_assert(sizeof(BatchGraph) == 80, "BatchGraph layout");

struct GenArgs {
    // single graph
    int32_t *G;               // targets, row-major [n][d]
    uint64_t n;
    uint64_t T;
    uint32_t key0, key1;
    // batch: chunks are handed out round by round -- chunk r of every graph
    // that has more than r chunks, then chunk r+1, ... -- so each graph's
    // chunks are claimed in ascending order while many graphs progress side
    // by side. Rounds with the same set of active graphs form a segment:
    // segment s starts at global chunk seg_base[s] and round seg_round[s], and
    // its rounds cover graphs order[0 .. seg_active[s]) (graphs sorted by
    // chunk count, descending). seg_base has nseg + 1 entries.
    const BatchGraph *graphs;
    const uint32_t *order;
    const uint64_t *seg_base;
    const uint64_t *seg_round;
    const uint64_t *seg_active;
    uint64_t nseg;
    // common
    uint32_t chunk;           // work claimed per wavefront claim (multiple of 64; d > kNarrowMaxD:
                              // nodes, wide_geometry(d).claim)
    uint32_t early_chunk;     // claim size below early_nodes (single graph)
    uint64_t early_nodes;     // below this node index every slot is published as accepted
    uint64_t watchdog_ticks;  // wall-clock ticks a wavefront may run
    DeviceScratch *scratch;
    uint32_t sleep_cap;       // longest wavefront back-off, in 256-clock sleeps (>= 1)
    uint32_t wide_mask;       // later pending slots are re-read when (pass & wide_mask) == 0
    // batch: finished nodes per graph (device), and per-graph flags in mapped
    // host memory set to 1 once a graph is complete and published (nullable)
    uint32_t *done_nodes;
    uint32_t *done_flag;
};

// The first nodes of a graph, generated in one workgroup's LDS: a single
// graph (graphs == nullptr), or graph pidx[blockIdx.x] of a batch, whose
// nodes [d+1, start) are generated.
struct PrefixArgs {
    int32_t *G;               // targets, row-major [n][d]
    uint64_t T;
    uint32_t key0, key1;
    uint32_t model;           // PPA_MODEL_*: which seed rows
    uint32_t d;               // set by the launchers (batch: per graph)
    uint64_t prefix;          // nodes [d+1, prefix) are generated; prefix * d <= kPrefixSlots
    uint64_t watchdog_ticks;
    DeviceScratch *scratch;
    const BatchGraph *graphs;
    const uint32_t *pidx;
    uint32_t *done_flag;      // batch: per-graph host flag, raised here for a graph the prefix completes
};

// int32 slots of the LDS-resident prefix, over the kPrefixLanes lanes of its
// one workgroup. A single graph of d >= kWidePrefixMinD (16-lane groups) may
// take the workgroup's whole LDS: kPrefixSlotsWide slots and the 4-byte abort
// word are 160,004 of gfx950's 163,840 bytes. Other prefixes, and every
// prefix of a batch (whose workgroups then share a CU two by two), stay in
// 64 KB: kPrefixSlots and the abort word are 65,284 bytes.
constexpr uint32_t kPrefixSlots = 16320;
constexpr uint32_t kPrefixSlotsWide = 40000;
constexpr uint32_t kWidePrefixMinD = 9;
constexpr uint32_t kPrefixMaxD = 32;  // a node's slots fit one half-wavefront group
constexpr uint32_t kPrefixLanes = 1024;
inline uint32_t prefix_slots(uint32_t d, bool wide)
{
    return wide && d >= kWidePrefixMinD ? kPrefixSlotsWide : kPrefixSlots;
}
inline uint64_t prefix_nodes(uint64_t n, uint32_t d, bool wide)
{
    const uint64_t p = prefix_slots(d, wide) / d;
    return n < p ? n : p;
}
// First node the persistent generator claims for a graph: the end of its
// LDS-resident prefix, or d + 1 without one. wide: a single graph (the
// 160 KB prefix is allowed). prefix_cap is the tuning's prefix_nodes:
// 1 = none, N = at most N, 0 = the default:
// * none where a node makes at most 4 copy draws on average, d(1-p) <= 4
//   (one CU would walk nodes the full grid takes in parallel: 5-15 % slower);
// * as many nodes as LDS holds where a wide prefix is allowed and a node
//   makes at least 8 copy draws (d = 16: 2,500 nodes took 4-9 % more off
//   calls of n = 2e4 .. 1e7 than 1,041 nodes did);
// * else 1024 + d + 1 nodes, one per lane (at d = 8 longer prefixes cost
//   1-10 %). Above 4 copies per node the prefix takes 25-60 % off a call
//   (d = 16, p = 0.1, n = 1e5: 4.0 -> 1.8 ms).
inline uint64_t prefix_start(uint64_t n, uint32_t d, uint64_t T, uint32_t prefix_cap, bool wide)
{
    if (prefix_cap == 1 || d > kPrefixMaxD) return (uint64_t)d + 1;
    uint64_t p = prefix_nodes(n, d, wide);
    if (prefix_cap != 0) {
        p = p < prefix_cap ? p : prefix_cap;
    } else {
        const double copies = (double)d * (1.0 - (double)T / 4294967296.0);
        if (copies <= 4.0) return (uint64_t)d + 1;
        if (!(wide && d >= kWidePrefixMinD && copies >= 8.0)) {
            const uint64_t one_per_lane = (uint64_t)kPrefixLanes + d + 1;
            p = p < one_per_lane ? p : one_per_lane;
        }
    }
    return p > (uint64_t)d + 1 ? p : (uint64_t)d + 1;
}

struct FillArgs {
    int32_t *G;
    int64_t *row_ptr;  // nullable
    uint64_t n;
    uint32_t d;
    uint32_t model;    // PPA_MODEL_*: which seed rows
};

struct BatchFillArgs {
    const BatchGraph *graphs;
    const uint64_t *slot_base;  // exclusive prefix of n_i * d_i, count + 1 entries
    const uint64_t *row_base;   // exclusive prefix of n_i + 1, count + 1 entries
    uint64_t count;
    uint32_t *done_nodes;       // per-graph finished-node counters, zeroed here
};

hipError_t launch_reset_scratch(DeviceScratch *s, uint64_t first, hipStream_t stream);
hipError_t launch_fill(const FillArgs &a, hipStream_t stream);
hipError_t launch_fill_batch(const BatchFillArgs &a, uint64_t total_slots, uint64_t total_rows,
                             hipStream_t stream);
hipError_t launch_generate(uint32_t d, const GenArgs &a, uint32_t blocks, uint32_t block_threads,
                           hipStream_t stream);
// Nodes d+1 .. a.prefix-1 of a single graph whose fill has run; one workgroup.
hipError_t launch_prefix(uint32_t d, const PrefixArgs &a, hipStream_t stream);
// The prefixes of `count` graphs of a batch (a.graphs[a.pidx[i]]); one workgroup each.
hipError_t launch_prefix_batch(const PrefixArgs &a, uint32_t count, hipStream_t stream);
hipError_t launch_generate_batch(const GenArgs &a, uint32_t blocks, uint32_t block_threads, hipStream_t stream);
hipError_t generate_occupancy(uint32_t d, uint32_t block_threads, int *blocks_per_cu);
hipError_t launch_degree_atomic(const int32_t *G, uint64_t n, uint32_t d, uint32_t *deg, int32_t *status,
                                hipStream_t stream);
// Adjusts the seed nodes' degrees of a non-copy model (one workgroup).
hipError_t launch_seed_degree(uint32_t *deg, uint32_t d, uint32_t model, hipStream_t stream);
hipError_t launch_degree_histogram(const uint32_t *deg, uint64_t n, uint64_t *hist, uint64_t hist_len,
                                   DeviceScratch *s, hipStream_t stream);

constexpr uint32_t kBatchDMax = 16;  // d of 0 selects the batch kernel in generate_occupancy

// Out-degrees above kNarrowMaxD run the wide generator: 64-thread workgroups
// (one wavefront) whose streams each work on one 64-slot window of a node;
// each node in flight keeps its accepted targets in an LDS hash set of
// 1 << hbits int32 words, the smallest power of two >= 2d (at most half full;
// 64 KB at d = 8192).
constexpr uint32_t kNarrowMaxD = 32;
constexpr uint32_t kWideThreads = 64;
constexpr uint32_t kWideWinWords = 256;  // a stream's window table: 128 keys, 128 lowest lanes
inline uint32_t wide_hash_bits(uint32_t d)
{
    uint32_t h = 1;
    while ((1u << h) < 2u * d) ++h;
    return h;
}
struct WideGeometry {
    uint32_t streams;  // windows in flight per wavefront: 8, 4 or 1
    uint32_t srows;    // attempts a rejected slot draws at once, at most: rows of 64 words per
                       // stream (64 with one stream: the first nodes' retry tails grow with d,
                       // the last slot of node d + 1 takes about d / 2 attempts)
    uint32_t prows;    // more rows for the gate window of the wavefront's oldest node, whose
                       // slots are on the critical path (kb up to srows + prows there)
    uint32_t nslots;   // nodes in flight per wavefront (a hash set each): one more than a claim,
                       // up to 8, within 32 KB of sets
    uint32_t hbits;
    uint32_t claim;    // nodes per claim (GenArgs::chunk)
};
// Streams and claim sizes by out-degree, as measured on n*d = 2^30 .. 2^31
// slots at p = 0.2: more streams overlap more windows, but every window
// behind a node's gate re-reads its still-empty sources while it waits, which
// past d = 192 costs more than the overlap gains (d = 256: 110 ms with one
// stream, 124 with 4, 172 with 8; d = 1024: 264 / 353 / 588 / 1102 ms with
// 1 / 2 / 4 / 8). Claims cost one atomic on one counter, which sustains about
// 7e7 a second: at d = 33, 16 nodes a claim ran in 59 ms, 8 in 68, 2 in 136;
// larger claims queue nodes behind each other (d = 64: 49 ms at 8, 92 at 32).
inline WideGeometry wide_geometry(uint32_t d)
{
    WideGeometry g;
    g.hbits = wide_hash_bits(d);
    if (d > 192) {
        // one window at a time, one node per wavefront: 64 rows for any slot.
        // Fewer rows fit more workgroups per CU but ran slower: d = 256 at
        // n = 2^22 107.4 ms with 64 rows (8 per CU), 110.6 with 32, 114.3 with
        // 16 (22 per CU), 115.9 with 8; d = 1024 262.5 / 268.7 / 272.6 / 278.9
        g.streams = 1u;
        g.srows = 64u;
        g.prows = 0u;
        g.claim = 1u;
        g.nslots = 1u;
        return g;
    }
    g.streams = d <= 96 ? 8u : 4u;
    g.claim = d <= 40 ? 16u : (d <= 64 ? 8u : (d <= 128 ? 4u : 2u));
    // LDS bounds the wavefronts per CU (one per 64-thread workgroup), and the
    // stream rows are much of it: 4 rows a stream held 8 workgroups per CU at
    // d = 33 (20 KB each) and 6 at d = 96 (24 KB); 1 row at d <= 40 and 2
    // above fit 11 and 8: d = 33 57.7 -> 51.1 ms (53.9 with 2 rows), d = 96
    // 70.6 -> 61.0, d = 48, 64 and 128 within 1 % (n * d = 2^30 .. 2^31).
    g.srows = d <= 40 ? 1u : 2u;
    g.prows = d > 64 ? 12u : 0u;
    const uint32_t by_lds = 8192u >> g.hbits;  // hbits <= 9 here
    const uint32_t want = g.claim + 1u < 8u ? g.claim + 1u : 8u;
    g.nslots = want < by_lds ? want : by_lds;
    return g;
}
inline size_t wide_lds_bytes(const WideGeometry &g)
{
    return (((size_t)g.nslots << g.hbits) + (size_t)g.streams * kWideWinWords +
            ((size_t)g.streams * g.srows + g.prows) * 64u) *
           4u;
}

}  // namespace ppa
