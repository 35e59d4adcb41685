// ppa_degree.hip -- partitioned degree pass for gfx950.
//
// Replaces the host loops of cuPPA::printDegreeDistribution (cuPPA.cu:115-160)
// and PAHelper::outputSCMDegreeDistribution (PAHelper.hpp:15-67):
// deg(v) = d + #{slots of nodes t >= d+1 naming v}.
//
// One device-scope atomic per slot is bound by the memory-side atomic rate
// (a few 1e9 per second on this part), so the slots are instead partitioned
// by target into "leaves" of 2^14 consecutive nodes and each leaf is counted
// in LDS (64 KB of counters per workgroup):
//   * level 1 partitions all slots by v >> 22 (up to 512 bins), level 2 each
//     level-1 bucket by bits 14..21 (256 bins); graphs of at most 2^22 nodes
//     need one level (v >> 14), and of at most 2^14 nodes none;
//   * a partition pass counts bins per workgroup (LDS histogram), scans the
//     counts in (segment, bin, workgroup) order, and scatters tile by tile:
//     each 4096-slot tile is ordered by bin in LDS first, so each bin's run
//     leaves the workgroup as contiguous stores;
//   * a leaf of at most kPiece slots is counted by one workgroup, which
//     stores its 2^14 degrees with plain coalesced stores; larger leaves (the
//     hubs of a preferential-attachment graph) are split into pieces whose
//     LDS counts are merged with one atomic per node and piece.
// Slot values outside [0, n) (an incomplete graph) are skipped everywhere.
#include "ppa_degree.hpp"
#include "ppa_device.hpp"

#include <algorithm>

namespace ppa {
namespace {

constexpr int kThreads = 256;
constexpr int kPerThread = 16;
constexpr int kTile = kThreads * kPerThread;  // slots a count workgroup reads at once
constexpr int kScatterThreads = 512;           // scatter workgroups order 8192 slots at once
constexpr uint32_t kMaxBins = 512;
constexpr uint32_t kLeafBits = 14;
constexpr uint32_t kLeafNodes = 1u << kLeafBits;
constexpr uint32_t kLevel1Shift = 22;
constexpr uint32_t kLevel2Bins = 256;
constexpr uint64_t kPiece = 1ull << 20;       // slots of one leaf-count work item
constexpr uint32_t kScanChunk = 4096;         // entries per scan workgroup
constexpr uint32_t kScanParts = 256;          // scan workgroups at most (one top-level pass)
constexpr uint32_t kMaxBlocks = 2048;         // partition workgroups at most

struct Geometry {
    uint64_t first = 0, slots = 0;  // generated slots [first, first + slots) of G
    uint32_t levels = 0;            // partition passes: 0, 1 or 2
    uint32_t bins1 = 0;             // bins of the first pass
    uint32_t leaves = 1;            // leaf entries (Lcap)
    uint32_t blocks1 = 0, blocks2 = 0;
    uint64_t counts = 0;            // entries of the counts array
};

Geometry geometry(uint64_t n, uint32_t d, uint32_t num_cus)
{
    Geometry g;
    g.first = (uint64_t)(d + 1) * d;
    g.slots = n > (uint64_t)d + 1 ? (n - d - 1) * (uint64_t)d : 0;
    const uint64_t leaves = (n + kLeafNodes - 1) >> kLeafBits;
    // The count and scatter kernels of a pass share the workgroup ranges. 6
    // per CU: the 512-thread scatter (4 resident per CU) then runs a second,
    // partly empty round, yet 4 per CU (one round) measured the same 71.8 ms
    // at n = 2^30, d = 8: the scatter is bound by memory, not by residency.
    uint64_t b = std::min<uint64_t>((uint64_t)num_cus * 6, kMaxBlocks);
    b = std::max<uint64_t>(1, std::min<uint64_t>(b, (g.slots + 65535) / 65536));
    g.blocks1 = (uint32_t)b;
    if (leaves <= 1) {
        g.levels = 0;
        g.leaves = 1;
    } else if (leaves <= kLevel2Bins) {
        g.levels = 1;
        g.bins1 = (uint32_t)leaves;
        g.leaves = g.bins1;
        g.counts = (uint64_t)g.bins1 * g.blocks1 + 1;
    } else {
        g.levels = 2;
        g.bins1 = (uint32_t)((n + (1ull << kLevel1Shift) - 1) >> kLevel1Shift);
        g.leaves = g.bins1 * kLevel2Bins;
        g.blocks2 = std::max<uint32_t>(g.blocks1, 2 * g.bins1 + 1);
        g.counts = std::max<uint64_t>((uint64_t)g.bins1 * g.blocks1, (uint64_t)kLevel2Bins * g.blocks2) + 1;
    }
    return g;
}

// One partition pass: segments of the input, each split into bins by
// (v >> shift) & mask over its own workgroups.
struct PartPlan {
    const uint32_t *seg_b0;     // nseg + 1 entries: first workgroup of each segment; [nseg] = workgroups used
    const uint64_t *seg_start;  // nseg + 1 entries: first slot of each segment; [nseg] = end
    uint32_t nseg, nbins, shift, mask;
};

struct BlockRange {
    uint64_t lo, hi;      // slots [lo, hi) of the input
    uint64_t count_base;  // counts index of (segment, bin 0, workgroup 0)
    uint32_t nb;          // workgroups of the segment
    uint32_t rank;        // this workgroup's rank in it
    uint32_t seg;
};

__device__ BlockRange block_range(const PartPlan &pl, uint32_t b)
{
    // the last segment s < nseg with seg_b0[s] <= b: empty segments share
    // their successor's seg_b0, so this is the non-empty one
    uint32_t lo = 0, hi = pl.nseg;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (pl.seg_b0[mid] <= b) lo = mid;
        else hi = mid;
    }
    BlockRange r;
    r.seg = lo;
    r.nb = pl.seg_b0[lo + 1] - pl.seg_b0[lo];
    r.rank = b - pl.seg_b0[lo];
    const uint64_t start = pl.seg_start[lo], end = pl.seg_start[lo + 1];
    const uint64_t per = (end - start + r.nb - 1) / r.nb;
    r.lo = start + (uint64_t)r.rank * per;
    if (r.lo > end) r.lo = end;
    r.hi = r.lo + per < end ? r.lo + per : end;
    r.count_base = (uint64_t)pl.nbins * pl.seg_b0[lo];
    return r;
}

__device__ __forceinline__ bool valid(int32_t v, uint64_t n) { return v >= 0 && (uint64_t)v < n; }

__device__ __forceinline__ uint64_t shfl_up_u64(uint64_t x, int off)
{
    const uint32_t lo = __shfl_up((uint32_t)x, off, kWave);
    const uint32_t hi = __shfl_up((uint32_t)(x >> 32), off, kWave);
    return ((uint64_t)hi << 32) | lo;
}

// Exclusive scan of one value per thread over a workgroup of NT threads;
// every thread must call it. red holds NT / 64 entries.
template <int NT>
__device__ uint64_t block_exclusive_scan(uint64_t v, uint64_t *total, uint64_t *red)
{
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint64_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint64_t y = shfl_up_u64(x, off);
        if ((int)lane >= off) x += y;
    }
    if (lane == 63u) red[wave] = x;
    __syncthreads();
    uint64_t before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) {
        const uint64_t s = red[w];
        if ((uint32_t)w < wave) before += s;
        all += s;
    }
    __syncthreads();
    *total = all;
    return before + x - v;
}

__global__ void fill_u64_kernel(uint64_t *p, uint64_t len, uint64_t value)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += stride) p[i] = value;
}

__global__ void fill_deg_kernel(uint32_t *deg, uint64_t n, uint32_t d)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) deg[v] = d;
}

// A single segment over all slots, or a single leaf holding them.
__global__ void plan_whole_kernel(uint32_t *seg_b0, uint64_t *seg_start, uint32_t blocks, uint64_t slots,
                                  uint64_t *leaf_size)
{
    if (threadIdx.x == 0) {
        seg_b0[0] = 0;
        seg_b0[1] = blocks;
        seg_start[0] = 0;
        seg_start[1] = slots;
        if (leaf_size) leaf_size[0] = slots;
    }
}

// Level-2 segments = level-1 buckets; workgroups in proportion to their size.
__global__ __launch_bounds__(kMaxBins) void plan_level2_kernel(const uint64_t *offsets1, uint32_t blocks1,
                                                               uint32_t nseg, uint32_t blocks2, uint32_t *seg_b0,
                                                               uint64_t *seg_start)
{
    __shared__ uint64_t start[kMaxBins + 1];
    __shared__ uint64_t red[kMaxBins / 64];
    const uint32_t s = threadIdx.x;
    // bucket s begins at the scanned count of (bin s, workgroup 0); the
    // entry after the last bucket holds the total
    if (s < nseg) start[s] = offsets1[(uint64_t)s * blocks1];
    if (s == 0) start[nseg] = offsets1[(uint64_t)nseg * blocks1];
    __syncthreads();
    const uint64_t total = start[nseg];
    const uint64_t per = total ? (total + (blocks2 - nseg) - 1) / (blocks2 - nseg) : 1;
    const uint64_t size = s < nseg ? start[s + 1] - start[s] : 0;
    const uint64_t nb = size ? (size + per - 1) / per : 0;  // sum <= (blocks2 - nseg) + nseg
    uint64_t used = 0;
    const uint64_t b0 = block_exclusive_scan<kMaxBins>(nb, &used, red);
    if (s < nseg) {
        seg_b0[s] = (uint32_t)b0;
        seg_start[s] = start[s];
    }
    if (s == 0) {
        seg_b0[nseg] = (uint32_t)used;
        seg_start[nseg] = total;
    }
}

// Bin 0 of a partition holds the smallest node ids, which in a
// preferential-attachment graph are named by far the most slots (a third of
// them at the first level): its updates are counted per lane, or ranked per
// wavefront, instead of contending for one LDS word.

// counts[(segment, bin, workgroup)] = slots of the workgroup's range in bin;
// with leaf_size, also the slots of every (segment, bin) leaf.
__global__ __launch_bounds__(kThreads) void part_count_kernel(const int32_t *in, PartPlan pl, uint64_t n,
                                                              uint64_t *counts, uint64_t *leaf_size)
{
    __shared__ uint32_t hist[kMaxBins];
    if (blockIdx.x >= pl.seg_b0[pl.nseg]) return;
    const BlockRange r = block_range(pl, blockIdx.x);
    for (uint32_t i = threadIdx.x; i < pl.nbins; i += kThreads) hist[i] = 0;
    __syncthreads();
    // kPerThread reads in flight per thread; the next tile's are issued
    // before this tile's LDS updates
    uint32_t hot = 0;  // this lane's slots in bin 0
    int32_t v[kPerThread];
#pragma unroll
    for (int j = 0; j < kPerThread; ++j) {
        const uint64_t i = r.lo + (uint64_t)j * kThreads + threadIdx.x;
        v[j] = i < r.hi ? in[i] : -1;
    }
    for (uint64_t t0 = r.lo; t0 < r.hi; t0 += kTile) {
        int32_t nv[kPerThread];
#pragma unroll
        for (int j = 0; j < kPerThread; ++j) {
            const uint64_t i = t0 + kTile + (uint64_t)j * kThreads + threadIdx.x;
            nv[j] = i < r.hi ? in[i] : -1;
        }
#pragma unroll
        for (int j = 0; j < kPerThread; ++j) {
            if (valid(v[j], n)) {
                const uint32_t bin = ((uint32_t)v[j] >> pl.shift) & pl.mask;
                if (bin == 0) ++hot;
                else atomicAdd(&hist[bin], 1u);
            }
        }
#pragma unroll
        for (int j = 0; j < kPerThread; ++j) v[j] = nv[j];
    }
    hot = wave_sum(hot);
    if ((threadIdx.x & 63u) == 0 && hot) atomicAdd(&hist[0], hot);
    __syncthreads();
    for (uint32_t bin = threadIdx.x; bin < pl.nbins; bin += kThreads) {
        const uint32_t c = hist[bin];
        counts[r.count_base + (uint64_t)bin * r.nb + r.rank] = c;
        if (leaf_size && c)
            atomicAdd((unsigned long long *)&leaf_size[(uint64_t)r.seg * pl.nbins + bin], (unsigned long long)c);
    }
}

// Exclusive scan of counts[0, len) into offsets[0, len], offsets[len] = total:
// workgroup sums, one top-level scan, then per-chunk scans.
__global__ __launch_bounds__(kThreads) void scan_reduce_kernel(const uint64_t *counts, uint64_t len,
                                                               uint64_t *partials)
{
    __shared__ uint64_t red[kThreads / 64];
    const uint64_t base = (uint64_t)blockIdx.x * kScanChunk;
    uint64_t s = 0;
    for (uint32_t i = threadIdx.x; i < kScanChunk; i += kThreads)
        if (base + i < len) s += counts[base + i];
    uint64_t total = 0;
    (void)block_exclusive_scan<kThreads>(s, &total, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanParts) void scan_top_kernel(uint64_t *partials, uint32_t nparts,
                                                              uint64_t *total_out)
{
    __shared__ uint64_t red[kScanParts / 64];
    const uint64_t v = threadIdx.x < nparts ? partials[threadIdx.x] : 0;
    uint64_t total = 0;
    const uint64_t ex = block_exclusive_scan<kScanParts>(v, &total, red);
    if (threadIdx.x < nparts) partials[threadIdx.x] = ex;
    if (threadIdx.x == 0) *total_out = total;
}

__global__ __launch_bounds__(kThreads) void scan_apply_kernel(const uint64_t *counts, uint64_t len,
                                                              const uint64_t *partials, uint64_t *offsets)
{
    __shared__ uint64_t red[kThreads / 64];
    constexpr uint32_t kPer = kScanChunk / kThreads;  // consecutive entries per thread
    const uint64_t base = (uint64_t)blockIdx.x * kScanChunk + (uint64_t)threadIdx.x * kPer;
    uint64_t v[kPer];
    uint64_t s = 0;
#pragma unroll
    for (uint32_t j = 0; j < kPer; ++j) {
        v[j] = base + j < len ? counts[base + j] : 0;
        s += v[j];
    }
    uint64_t total = 0;
    uint64_t run = block_exclusive_scan<kThreads>(s, &total, red) + partials[blockIdx.x];
#pragma unroll
    for (uint32_t j = 0; j < kPer; ++j) {
        if (base + j < len) offsets[base + j] = run;
        run += v[j];
    }
}

// Moves every valid slot of the workgroup's range to its bin's run of the
// output; tile by tile, slots are grouped by bin in LDS before the stores.
// TH threads place TH * kPerThread slots per tile: at 512 threads a bin's run
// in a tile averages 32 slots at 256 bins, so most of a tile's stores fill
// whole 64-byte lines.
template <int TH>
__global__ __launch_bounds__(TH) void part_scatter_kernel(const int32_t *in, int32_t *out, PartPlan pl, uint64_t n,
                                                          const uint64_t *offsets)
{
    constexpr int kT = TH * kPerThread;
    constexpr uint32_t kBinsPerThread = (kMaxBins + TH - 1) / TH;
    __shared__ uint64_t cur[kMaxBins];      // next output slot of each bin
    __shared__ uint32_t thist[kMaxBins];    // slots of each bin in this tile
    __shared__ uint32_t tstart[kMaxBins];   // their exclusive scan
    __shared__ int32_t tile[kT];
    __shared__ uint64_t red[TH / 64];
    if (blockIdx.x >= pl.seg_b0[pl.nseg]) return;
    const BlockRange r = block_range(pl, blockIdx.x);
    for (uint32_t bin = threadIdx.x; bin < pl.nbins; bin += TH) {
        cur[bin] = offsets[r.count_base + (uint64_t)bin * r.nb + r.rank];
        thist[bin] = 0;
    }
    __syncthreads();
    int32_t v[kPerThread];
#pragma unroll
    for (int j = 0; j < kPerThread; ++j) {
        const uint64_t i = r.lo + (uint64_t)j * TH + threadIdx.x;
        v[j] = i < r.hi ? in[i] : -1;
    }
    for (uint64_t t0 = r.lo; t0 < r.hi; t0 += kT) {
        uint32_t rank[kPerThread];
#pragma unroll
        for (int j = 0; j < kPerThread; ++j) {
            const bool ok = valid(v[j], n);
            const uint32_t bin = ((uint32_t)v[j] >> pl.shift) & pl.mask;
            // bin 0: one LDS atomic per wavefront, ranks by lane order
            const uint64_t m0 = __ballot(ok && bin == 0);
            uint32_t base0 = 0;
            if (m0 != 0) {
                const uint32_t leader = (uint32_t)__builtin_ctzll(m0);
                if ((threadIdx.x & 63u) == leader) base0 = atomicAdd(&thist[0], (uint32_t)__popcll(m0));
                base0 = (uint32_t)__builtin_amdgcn_readlane((int)base0, (int)leader);
            }
            rank[j] = !ok ? 0u : bin == 0 ? base0 + rank_below(m0) : atomicAdd(&thist[bin], 1u);
        }
        __syncthreads();
        // bins [kBinsPerThread * tid, +kBinsPerThread) per thread
        uint32_t c[kBinsPerThread], csum = 0;
#pragma unroll
        for (uint32_t q = 0; q < kBinsPerThread; ++q) {
            const uint32_t b = kBinsPerThread * threadIdx.x + q;
            c[q] = b < pl.nbins ? thist[b] : 0u;
            csum += c[q];
        }
        uint64_t tot = 0;
        uint32_t ex = (uint32_t)block_exclusive_scan<TH>(csum, &tot, red);
#pragma unroll
        for (uint32_t q = 0; q < kBinsPerThread; ++q) {
            const uint32_t b = kBinsPerThread * threadIdx.x + q;
            if (b < pl.nbins) tstart[b] = ex;
            ex += c[q];
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kPerThread; ++j)
            if (valid(v[j], n)) tile[tstart[((uint32_t)v[j] >> pl.shift) & pl.mask] + rank[j]] = v[j];
        // the next tile's reads are in flight while this one is stored
#pragma unroll
        for (int j = 0; j < kPerThread; ++j) {
            const uint64_t i = t0 + kT + (uint64_t)j * TH + threadIdx.x;
            v[j] = i < r.hi ? in[i] : -1;
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < (uint32_t)tot; i += TH) {
            const int32_t w = tile[i];
            const uint32_t bin = ((uint32_t)w >> pl.shift) & pl.mask;
            out[cur[bin] + (i - tstart[bin])] = w;
        }
        __syncthreads();
        for (uint32_t bin = threadIdx.x; bin < pl.nbins; bin += TH) {
            cur[bin] += thist[bin];
            thist[bin] = 0;
        }
        __syncthreads();
    }
}

// leaf_start = exclusive scan of leaf_size; piece_base = exclusive scan of
// max(1, ceil(leaf_size / kPiece)); one workgroup.
constexpr int kPlanThreads = 1024;

__global__ __launch_bounds__(kPlanThreads) void leaf_plan_kernel(const uint64_t *leaf_size, uint32_t leaves,
                                                                 uint64_t *leaf_start, uint64_t *piece_base)
{
    __shared__ uint64_t red[kPlanThreads / 64];
    const uint32_t per = (leaves + kPlanThreads - 1) / kPlanThreads;
    const uint32_t lo = threadIdx.x * per, hi = lo + per < leaves ? lo + per : leaves;
    uint64_t s = 0, p = 0;
    for (uint32_t l = lo; l < hi; ++l) {
        const uint64_t z = leaf_size[l];
        s += z;
        p += z > kPiece ? (z + kPiece - 1) / kPiece : 1;
    }
    uint64_t stot = 0, ptot = 0;
    uint64_t srun = block_exclusive_scan<kPlanThreads>(s, &stot, red);
    uint64_t prun = block_exclusive_scan<kPlanThreads>(p, &ptot, red);
    for (uint32_t l = lo; l < hi; ++l) {
        const uint64_t z = leaf_size[l];
        leaf_start[l] = srun;
        piece_base[l] = prun;
        srun += z;
        prun += z > kPiece ? (z + kPiece - 1) / kPiece : 1;
    }
    if (threadIdx.x == 0) {
        leaf_start[leaves] = stot;
        piece_base[leaves] = ptot;
    }
}

// deg of every node of a leaf from LDS counters: a one-piece leaf stores
// d + count, a split leaf adds its piece's counts to deg (set to d before).
__global__ __launch_bounds__(kThreads) void leaf_count_kernel(const int32_t *in, uint64_t n, uint32_t d,
                                                              uint32_t leaves, const uint64_t *leaf_start,
                                                              const uint64_t *piece_base, uint32_t *deg)
{
    __shared__ uint32_t cnt[kLeafNodes];
    const uint64_t pieces = piece_base[leaves];
    for (uint64_t q = blockIdx.x; q < pieces; q += gridDim.x) {
        uint32_t lo = 0, hi = leaves;  // the leaf: last l with piece_base[l] <= q
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (piece_base[mid] <= q) lo = mid;
            else hi = mid;
        }
        const uint64_t np = piece_base[lo + 1] - piece_base[lo];
        const uint64_t first = leaf_start[lo] + (q - piece_base[lo]) * kPiece;
        const uint64_t end = first + kPiece < leaf_start[lo + 1] ? first + kPiece : leaf_start[lo + 1];
        for (uint32_t i = threadIdx.x; i < kLeafNodes; i += kThreads) cnt[i] = 0;
        __syncthreads();
        for (uint64_t t0 = first; t0 < end; t0 += kTile) {
            int32_t v[kPerThread];
#pragma unroll
            for (int j = 0; j < kPerThread; ++j) {
                const uint64_t i = t0 + (uint64_t)j * kThreads + threadIdx.x;
                v[j] = i < end ? in[i] : -1;
            }
#pragma unroll
            for (int j = 0; j < kPerThread; ++j)
                if (valid(v[j], n)) atomicAdd(&cnt[(uint32_t)v[j] & (kLeafNodes - 1)], 1u);
        }
        __syncthreads();
        const uint64_t base = (uint64_t)lo << kLeafBits;
        for (uint32_t i = threadIdx.x; i < kLeafNodes; i += kThreads) {
            const uint64_t v = base + i;
            if (v < n) {
                if (np == 1) deg[v] = d + cnt[i];
                else if (cnt[i]) atomicAdd(&deg[v], cnt[i]);
            }
        }
        __syncthreads();
    }
}

uint32_t grid_of(uint64_t work, uint32_t block, uint32_t cap)
{
    uint64_t g = (work + block - 1) / block;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(g, cap));
}

hipError_t scan(const uint64_t *counts, uint64_t len, const DegreeWork &w, hipStream_t stream)
{
    const uint32_t parts = (uint32_t)((len + kScanChunk - 1) / kScanChunk);
    if (parts > kScanParts) return hipErrorInvalidValue;
    hipLaunchKernelGGL(scan_reduce_kernel, dim3(parts), dim3(kThreads), 0, stream, counts, len, w.partials);
    hipLaunchKernelGGL(scan_top_kernel, dim3(1), dim3(kScanParts), 0, stream, w.partials, parts, w.offsets + len);
    hipLaunchKernelGGL(scan_apply_kernel, dim3(parts), dim3(kThreads), 0, stream, counts, len, w.partials,
                       w.offsets);
    return hipGetLastError();
}

// count, scan and scatter of one partition pass
hipError_t partition(const int32_t *in, int32_t *out, const PartPlan &pl, uint32_t grid, uint64_t n,
                     uint64_t count_len, uint64_t *leaf_size, const DegreeWork &w, hipStream_t stream)
{
    hipLaunchKernelGGL(part_count_kernel, dim3(grid), dim3(kThreads), 0, stream, in, pl, n, w.counts, leaf_size);
    hipError_t e = scan(w.counts, count_len, w, stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(part_scatter_kernel<kScatterThreads>, dim3(grid), dim3(kScatterThreads), 0, stream, in, out, pl, n,
                       (const uint64_t *)w.offsets);
    return hipGetLastError();
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// ---- in-edge CSR: a stable LSD radix sort of the slots by target ------------
// Pass p orders the (target, source) pairs by bits [8p, 8p+8) of the target,
// keeping the order of equal digits; after ceil(bits(n-1) / 8) passes the
// sources of each target are in slot order, which is ascending source order.
// Ranks are exact without LDS atomics: a wavefront owns 1024 consecutive
// slots of a tile (8192 slots for the 8 wavefronts of a workgroup) and walks
// them 64 at a time; lanes with equal digits are matched with 8 ballots and
// counted from a wavefront-private running count, and the wavefronts' counts
// are combined per digit. A digit's run in a tile averages 32 pairs, so most
// of its stores fill whole 64-byte lines.
constexpr uint32_t kRadixBits = 8;
constexpr uint32_t kRadixBins = 1u << kRadixBits;
constexpr uint32_t kLsdThreads = 512;
constexpr uint32_t kWaveShare = 1024;             // slots per wavefront per tile
constexpr uint32_t kSteps = kWaveShare / 64;      // 64-slot steps per wavefront

// Lanes among `act` whose digit equals this lane's; whole wavefront.
__device__ __forceinline__ uint64_t match_digit(uint32_t digit, bool act)
{
    uint64_t eq = __ballot(act);
#pragma unroll
    for (uint32_t b = 0; b < kRadixBits; ++b) {
        const bool bit = (digit >> b) & 1u;
        const uint64_t m = __ballot(bit);
        eq &= bit ? m : ~m;
    }
    return eq;
}

// The number of valid pairs left by the first pass (offsets[len]) becomes
// the single segment of every later pass.
__global__ void plan_total_kernel(uint32_t *seg_b0, uint64_t *seg_start, uint32_t blocks, const uint64_t *total)
{
    if (threadIdx.x == 0) {
        seg_b0[0] = 0;
        seg_b0[1] = blocks;
        seg_start[0] = 0;
        seg_start[1] = *total;
    }
}

// One stable pass. vin == nullptr: the sources are the slot indices' rows,
// (first + i) / d (first pass, reading the targets array itself). kout may be
// null (no later pass needs the keys).
template <uint32_t TH>
__global__ __launch_bounds__(TH) void lsd_scatter_kernel(const int32_t *kin, const int32_t *vin, int32_t *kout,
                                                         int32_t *vout, PartPlan pl, uint64_t n, uint32_t d,
                                                         double inv_d, uint64_t first, const uint64_t *offsets)
{
    constexpr uint32_t kWaves = TH / 64;
    constexpr uint32_t kT = kWaves * kWaveShare;
    static_assert(TH >= kRadixBins, "one thread per digit");
    __shared__ uint64_t cur[kRadixBins];            // next output position per digit
    __shared__ uint32_t wrun[kWaves][kRadixBins];   // per-wavefront counts, then their prefix over wavefronts
    __shared__ uint32_t tstart[kRadixBins];         // tile start of each digit
    __shared__ int32_t tk[kT], tv[kT];
    __shared__ uint64_t red[kWaves];
    if (blockIdx.x >= pl.seg_b0[pl.nseg]) return;
    const BlockRange r = block_range(pl, blockIdx.x);
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
    const bool digit_thread = threadIdx.x < kRadixBins;
    if (digit_thread) cur[threadIdx.x] = offsets[r.count_base + (uint64_t)threadIdx.x * r.nb + r.rank];
    __syncthreads();
    for (uint64_t t0 = r.lo; t0 < r.hi; t0 += kT) {
        for (uint32_t b = lane; b < kRadixBins; b += 64) wrun[wave][b] = 0;
        int32_t k[kSteps], v[kSteps];
        uint32_t rk[kSteps];
#pragma unroll
        for (uint32_t j = 0; j < kSteps; ++j) {
            const uint64_t i = t0 + wave * kWaveShare + j * 64 + lane;
            k[j] = i < r.hi ? kin[i] : -1;
            v[j] = vin ? (i < r.hi ? vin[i] : 0) : (int32_t)div_small(first + i, d, inv_d);
        }
#pragma unroll
        for (uint32_t j = 0; j < kSteps; ++j) {
            const bool act = valid(k[j], n);
            const uint32_t dig = ((uint32_t)k[j] >> pl.shift) & (kRadixBins - 1);
            const uint64_t eq = match_digit(dig, act);
            const uint32_t below = rank_below(eq);
            const uint32_t base = act ? wrun[wave][dig] : 0u;
            __builtin_amdgcn_wave_barrier();
            if (act && below == 0) wrun[wave][dig] = base + (uint32_t)__popcll(eq);
            __builtin_amdgcn_wave_barrier();
            rk[j] = base + below;
        }
        __syncthreads();
        // thread = digit: prefix of the wavefronts' counts, and the tile total
        uint32_t tot = 0;
        if (digit_thread) {
#pragma unroll
            for (uint32_t w = 0; w < kWaves; ++w) {
                const uint32_t c = wrun[w][threadIdx.x];
                wrun[w][threadIdx.x] = tot;
                tot += c;
            }
        }
        uint64_t total = 0;
        const uint32_t ts = (uint32_t)block_exclusive_scan<TH>(tot, &total, red);
        if (digit_thread) tstart[threadIdx.x] = ts;
        __syncthreads();
#pragma unroll
        for (uint32_t j = 0; j < kSteps; ++j) {
            if (valid(k[j], n)) {
                const uint32_t dig = ((uint32_t)k[j] >> pl.shift) & (kRadixBins - 1);
                const uint32_t pos = tstart[dig] + wrun[wave][dig] + rk[j];
                tk[pos] = k[j];
                tv[pos] = v[j];
            }
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < (uint32_t)total; i += TH) {
            const int32_t kk = tk[i];
            const uint32_t dig = ((uint32_t)kk >> pl.shift) & (kRadixBins - 1);
            const uint64_t o = cur[dig] + (i - tstart[dig]);
            if (kout) kout[o] = kk;
            vout[o] = tv[i];
        }
        __syncthreads();
        if (digit_thread) cur[threadIdx.x] += tot;
        __syncthreads();
    }
}

// in_ptr[v] = first position of target v in the sorted keys, for v in
// (keys[i-1], keys[i]]; in_ptr[n] = total. A thread takes four consecutive
// positions per step with one 16-byte load (keys is 256-byte aligned, so
// position 4g is 16-byte aligned), plus the key before them: one element per
// thread with two 4-byte loads ran at 1.8 TB/s.
__global__ void csr_ptr_kernel(const int32_t *keys, const uint64_t *total_p, uint64_t n, uint64_t cap,
                               int64_t *in_ptr)
{
    const uint64_t total = *total_p < cap ? *total_p : cap;  // positions 0..total
    const uint64_t groups = total / 4 + 1;                   // group g covers positions 4g .. 4g+3
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < groups; g += stride) {
        const uint64_t i0 = g * 4;
        int64_t k[5];  // k[q] = key at position i0 + q - 1 (-1 before the first), n at and after total
        k[0] = i0 == 0 ? -1 : (int64_t)keys[i0 - 1];
        if (i0 + 4 <= total) {
            const int4 w = *reinterpret_cast<const int4 *>(keys + i0);
            k[1] = w.x;
            k[2] = w.y;
            k[3] = w.z;
            k[4] = w.w;
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) k[q + 1] = i0 + q < total ? (int64_t)keys[i0 + q] : (int64_t)n;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint64_t i = i0 + q;
            if (i > total) break;
            for (int64_t v = k[q] + 1; v <= k[q + 1]; ++v) in_ptr[v] = (int64_t)i;
        }
    }
}

// Undirected lists from the out rows and the in-edge CSR: node v's list is
// its d targets (slot order), then the sources t >= d+1 of the slots naming it
// (ascending); ptr[v] = v*d + in_ptr[v] - d*min(v, d+1). The in-edge sort's
// position i of (v, t) maps to list index i + d*(v - d) for v > d; a seed node
// v <= d keeps position i, and its d seed-row sources (t <= d), which sort
// first, give their places to its own row.
__global__ void undirected_in_kernel(const int32_t *keys, const int32_t *src, uint64_t total, uint32_t d,
                                     int32_t *adj)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const int32_t t = src[i];
        if ((uint32_t)t <= d) continue;
        const uint64_t v = (uint32_t)keys[i];
        adj[i + (v > d ? (v - d) * d : 0)] = t;
    }
}

__global__ void undirected_out_kernel(const int32_t *G, const int64_t *in_ptr, uint64_t n, uint32_t d, int64_t *ptr,
                                      int32_t *adj)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v <= n; v += stride) {
        const uint64_t seeds = v < (uint64_t)d + 1 ? v : (uint64_t)d + 1;
        const int64_t base = (int64_t)(v * d) + in_ptr[v] - (int64_t)(seeds * d);
        ptr[v] = base;
        if (v < n)
            for (uint32_t k = 0; k < d; ++k) adj[base + k] = G[v * d + k];
    }
}

struct CsrGeometry {
    uint64_t slots = 0;
    uint32_t passes = 1, blocks = 1;
    uint64_t counts = 0;
};

CsrGeometry csr_geometry(uint64_t n, uint32_t d, uint32_t num_cus)
{
    CsrGeometry g;
    g.slots = n * (uint64_t)d;
    uint32_t bits = 1;
    while (bits < 63 && (1ull << bits) < n) ++bits;  // targets are < n <= 2^bits
    g.passes = (bits + kRadixBits - 1) / kRadixBits;
    uint64_t b = std::min<uint64_t>((uint64_t)num_cus * 4, kMaxBlocks);
    g.blocks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(b, (g.slots + 65535) / 65536));
    g.counts = (uint64_t)kRadixBins * g.blocks + 1;
    return g;
}

}  // namespace

size_t degree_work_bytes(uint64_t n, uint32_t d, uint32_t num_cus)
{
    const Geometry g = geometry(n, d, num_cus);
    size_t b = 0;
    if (g.levels >= 1) b += align256(g.slots * 4);
    if (g.levels == 2) b += align256(g.slots * 4);
    b += 2 * align256(g.counts * 8);
    b += align256(kScanParts * 8);
    b += align256(2 * (kMaxBins + 1) * 4) + align256(2 * (kMaxBins + 1) * 8);
    b += 3 * align256(((size_t)g.leaves + 1) * 8);
    return b;
}

void degree_work_layout(void *base, uint64_t n, uint32_t d, uint32_t num_cus, DegreeWork *w)
{
    const Geometry g = geometry(n, d, num_cus);
    char *p = (char *)base;
    auto take = [&](size_t bytes) {
        char *q = p;
        p += align256(bytes);
        return (void *)q;
    };
    *w = DegreeWork{};
    if (g.levels >= 1) w->part_a = (int32_t *)take(g.slots * 4);
    if (g.levels == 2) w->part_b = (int32_t *)take(g.slots * 4);
    w->counts = (uint64_t *)take(g.counts * 8);
    w->offsets = (uint64_t *)take(g.counts * 8);
    w->partials = (uint64_t *)take(kScanParts * 8);
    w->seg_b0 = (uint32_t *)take(2 * (kMaxBins + 1) * 4);
    w->seg_start = (uint64_t *)take(2 * (kMaxBins + 1) * 8);
    w->leaf_size = (uint64_t *)take(((size_t)g.leaves + 1) * 8);
    w->leaf_start = (uint64_t *)take(((size_t)g.leaves + 1) * 8);
    w->piece_base = (uint64_t *)take(((size_t)g.leaves + 1) * 8);
    w->num_cus = num_cus;
}

hipError_t launch_degree_partitioned(const int32_t *G, uint64_t n, uint32_t d, uint32_t *deg, const DegreeWork &w,
                                     hipStream_t stream)
{
    const Geometry g = geometry(n, d, w.num_cus);
    hipLaunchKernelGGL(fill_deg_kernel, dim3(grid_of(n, kThreads, 8192)), dim3(kThreads), 0, stream, deg, n, d);
    if (g.slots == 0) return hipGetLastError();
    const int32_t *in = G + g.first;
    const int32_t *leaf_in = in;
    hipLaunchKernelGGL(fill_u64_kernel, dim3(grid_of(g.leaves, kThreads, 1024)), dim3(kThreads), 0, stream,
                       w.leaf_size, (uint64_t)g.leaves, 0ull);
    uint32_t *b0_1 = w.seg_b0, *b0_2 = w.seg_b0 + kMaxBins + 1;
    uint64_t *st_1 = w.seg_start, *st_2 = w.seg_start + kMaxBins + 1;
    hipError_t e = hipSuccess;
    if (g.levels == 0) {
        hipLaunchKernelGGL(plan_whole_kernel, dim3(1), dim3(64), 0, stream, b0_1, st_1, 1u, g.slots, w.leaf_size);
    } else if (g.levels == 1) {
        hipLaunchKernelGGL(plan_whole_kernel, dim3(1), dim3(64), 0, stream, b0_1, st_1, g.blocks1, g.slots,
                           (uint64_t *)nullptr);
        const PartPlan pl{b0_1, st_1, 1u, g.bins1, kLeafBits, 0xffu};
        e = partition(in, w.part_a, pl, g.blocks1, n, (uint64_t)g.bins1 * g.blocks1, w.leaf_size, w, stream);
        leaf_in = w.part_a;
    } else {
        hipLaunchKernelGGL(plan_whole_kernel, dim3(1), dim3(64), 0, stream, b0_1, st_1, g.blocks1, g.slots,
                           (uint64_t *)nullptr);
        const PartPlan p1{b0_1, st_1, 1u, g.bins1, kLevel1Shift, 0x1ffu};
        e = partition(in, w.part_a, p1, g.blocks1, n, (uint64_t)g.bins1 * g.blocks1, nullptr, w, stream);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(plan_level2_kernel, dim3(1), dim3(kMaxBins), 0, stream, (const uint64_t *)w.offsets,
                           g.blocks1, g.bins1, g.blocks2, b0_2, st_2);
        const uint64_t len2 = (uint64_t)kLevel2Bins * g.blocks2;
        hipLaunchKernelGGL(fill_u64_kernel, dim3(grid_of(len2, kThreads, 1024)), dim3(kThreads), 0, stream,
                           w.counts, len2, 0ull);
        const PartPlan p2{b0_2, st_2, g.bins1, kLevel2Bins, kLeafBits, 0xffu};
        e = partition(w.part_a, w.part_b, p2, g.blocks2, n, len2, w.leaf_size, w, stream);
        leaf_in = w.part_b;
    }
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(leaf_plan_kernel, dim3(1), dim3(kPlanThreads), 0, stream, (const uint64_t *)w.leaf_size,
                       g.leaves, w.leaf_start, w.piece_base);
    hipLaunchKernelGGL(leaf_count_kernel, dim3(std::max<uint32_t>(1, 2 * w.num_cus)), dim3(kThreads), 0, stream,
                       leaf_in, n, d, g.leaves, (const uint64_t *)w.leaf_start, (const uint64_t *)w.piece_base, deg);
    return hipGetLastError();
}

size_t csr_work_bytes(uint64_t n, uint32_t d, uint32_t num_cus)
{
    const CsrGeometry g = csr_geometry(n, d, num_cus);
    return 4 * align256(g.slots * 4) + 2 * align256(g.counts * 8) + align256(kScanParts * 8) + align256(2 * 4) +
           align256(2 * 8);
}

hipError_t launch_in_csr(const int32_t *G, uint64_t n, uint32_t d, int64_t *in_ptr, int32_t *in_src, void *work,
                         uint32_t num_cus, hipStream_t stream, const int32_t **sorted_keys)
{
    const CsrGeometry g = csr_geometry(n, d, num_cus);
    char *p = (char *)work;
    auto take = [&](size_t bytes) {
        char *q = p;
        p += align256(bytes);
        return (void *)q;
    };
    int32_t *ka = (int32_t *)take(g.slots * 4), *va = (int32_t *)take(g.slots * 4);
    int32_t *kb = (int32_t *)take(g.slots * 4), *vb = (int32_t *)take(g.slots * 4);
    DegreeWork w;
    w.counts = (uint64_t *)take(g.counts * 8);
    w.offsets = (uint64_t *)take(g.counts * 8);
    w.partials = (uint64_t *)take(kScanParts * 8);
    uint32_t *seg_b0 = (uint32_t *)take(2 * 4);
    uint64_t *seg_start = (uint64_t *)take(2 * 8);
    const uint64_t len = (uint64_t)kRadixBins * g.blocks;
    const double inv_d = 1.0 / d;
    hipError_t e = hipSuccess;
    hipLaunchKernelGGL(plan_whole_kernel, dim3(1), dim3(64), 0, stream, seg_b0, seg_start, g.blocks, g.slots,
                       (uint64_t *)nullptr);
    const int32_t *kin = G, *vin = nullptr;
    for (uint32_t pass = 0; pass < g.passes; ++pass) {
        const PartPlan pl{seg_b0, seg_start, 1u, kRadixBins, pass * kRadixBits, kRadixBins - 1};
        const bool last = pass + 1 == g.passes;
        int32_t *kout = (pass & 1) ? kb : ka;
        int32_t *vout = last ? in_src : ((pass & 1) ? vb : va);
        hipLaunchKernelGGL(part_count_kernel, dim3(g.blocks), dim3(kThreads), 0, stream, kin, pl, n, w.counts,
                           (uint64_t *)nullptr);
        if ((e = scan(w.counts, len, w, stream)) != hipSuccess) return e;
        hipLaunchKernelGGL(lsd_scatter_kernel<kLsdThreads>, dim3(g.blocks), dim3(kLsdThreads), 0, stream, kin, vin, kout,
                           vout, pl, n,
                           d, inv_d, 0ull, (const uint64_t *)w.offsets);
        if (pass == 0)  // later passes cover the valid pairs only
            hipLaunchKernelGGL(plan_total_kernel, dim3(1), dim3(64), 0, stream, seg_b0, seg_start, g.blocks,
                               (const uint64_t *)(w.offsets + len));
        kin = kout;
        vin = vout;
    }
    hipLaunchKernelGGL(csr_ptr_kernel, dim3(grid_of(g.slots + 1, kThreads, 16384)), dim3(kThreads), 0, stream, kin,
                       (const uint64_t *)(seg_start + 1), n, g.slots, in_ptr);
    if (sorted_keys) *sorted_keys = kin;
    return hipGetLastError();
}

hipError_t launch_undirected(const int32_t *G, uint64_t n, uint32_t d, const int64_t *in_ptr, const int32_t *keys,
                             const int32_t *in_src, int64_t *ptr, int32_t *adj, hipStream_t stream)
{
    hipLaunchKernelGGL(undirected_in_kernel, dim3(grid_of(n * d, kThreads, 16384)), dim3(kThreads), 0, stream, keys,
                       in_src, n * d, d, adj);
    hipLaunchKernelGGL(undirected_out_kernel, dim3(grid_of(n + 1, kThreads, 16384)), dim3(kThreads), 0, stream, G,
                       in_ptr, n, d, ptr, adj);
    return hipGetLastError();
}

}  // namespace ppa
