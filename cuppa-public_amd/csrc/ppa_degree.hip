// ppa_degree.hip -- partitioned degree pass and in-edge CSR for gfx950.
//
// Replaces the host loops of cuPPA::printDegreeDistribution (cuPPA.cu:115-160)
// and PAHelper::outputSCMDegreeDistribution (PAHelper.hpp:15-67):
// deg(v) = d + #{slots of nodes t >= d+1 naming v}.
//
// One device-scope atomic per slot is bound by the memory-side atomic rate
// (a few 1e9 per second on this part), so the slots are instead partitioned
// by target into "leaves" of 2^15 consecutive nodes and each leaf is counted
// in LDS (128 KB of counters per workgroup):
//   * graphs of more than 2^23 nodes take two partition levels: level 1 by
//     bucket v >> 23 (at most 256 buckets), level 2 each bucket by bits
//     15..22 (256 leaves per bucket); graphs of at most 2^23 nodes one level
//     (v >> 15), and of at most 2^15 nodes none;
//   * one count pass reads the slots once and counts, per workgroup, both
//     levels at once: the level-1 buckets and the (bucket, leaf) matrix, 128
//     buckets x 256 leaves of LDS counters per launch (one launch up to 2^30
//     nodes, two above). Scanned in (bucket, leaf, workgroup) order, the
//     matrix places level 2 without a second count pass: level-2 workgroup j
//     moves exactly the runs that level-1 workgroup j left in the buckets.
//     The same pass counts the 8192 hub nodes (the smallest ids) exactly and
//     leaves their slots out of the partitions;
//   * a scatter orders each 12288-slot tile by bin in LDS before storing it,
//     so each bin's run leaves the workgroup as contiguous stores; the last
//     level keeps only a key's low 15 bits (16-bit keys, half the bytes);
//   * a leaf of at most kPiece slots is counted by one workgroup, which
//     stores its 2^15 degrees with coalesced stores; larger leaves (the hubs
//     of a preferential-attachment graph) are split into pieces whose LDS
//     counts are merged with one atomic per node and piece.
// Slot values outside [0, n) (an incomplete graph) are skipped everywhere.
#include "ppa_degree.hpp"
#include "ppa_device.hpp"

#include <algorithm>

namespace ppa {
namespace {

constexpr int kThreads = 256;
constexpr int kPerThread = 16;
constexpr int kTile = kThreads * kPerThread;  // slots a count workgroup reads at once
constexpr uint32_t kMaxBins = 512;
constexpr uint32_t kScanChunk = 4096;         // entries per scan step of a workgroup
constexpr uint32_t kScanParts = 256;          // scan workgroups at most (one top-level pass)
constexpr uint32_t kMaxBlocks = 2048;         // partition workgroups at most

// degree pass
constexpr uint32_t kLeafBits = 15;            // a leaf: 2^15 nodes, 128 KB of LDS counters
constexpr uint32_t kLeafNodes = 1u << kLeafBits;
constexpr uint32_t kSubBits = 8;              // leaves per level-1 bucket of a two-level graph
constexpr uint32_t kSubBins = 1u << kSubBits;
constexpr uint32_t kBucketShift = kLeafBits + kSubBits;  // two-level bucket: v >> 23
constexpr uint32_t kMaxBuckets = 256;         // n <= 2^31
constexpr uint32_t kMatBuckets = 128;         // buckets one matrix count covers (128 KB of LDS)
constexpr uint32_t kHubNodes = 8192;          // nodes the count pass counts itself (32 KB of LDS)
constexpr int kCountThreads = 1024;           // count workgroups: one per CU by LDS, 16 waves
constexpr int kCountPer = 32;                 // slots per count thread and tile (16: 0.35 ms slower at 2^33 slots)
constexpr int kPartThreads = 512;             // scatter workgroups: 12288-slot tiles
constexpr int kPartPer = 24;                  // slots per scatter thread per tile (16, 28, 32: see below)
constexpr int kPartTile = kPartThreads * kPartPer;
constexpr uint32_t kPartPerCU = 4;            // partition workgroups per CU (8 measured 2 % slower;
                                              // 2: 46.0-46.4 ms against 46.2-48.5, inside the spread)
constexpr int kLeafThreads = 1024;
constexpr int kLeafUnroll = 4;                // 16-byte key loads in flight per leaf-count thread (8, 16: no faster;
                                              // a ballot per key to add equal keys at once: 4.3 -> 7.0 ms)
constexpr uint64_t kPiece = 1ull << 20;       // slots of one leaf-count work item
constexpr int kPlanThreads = 1024;

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

__global__ void fill_deg_kernel(uint32_t *deg, uint64_t n, uint32_t d)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) deg[v] = d;
}

// A single segment over all slots.
__global__ void plan_whole_kernel(uint32_t *seg_b0, uint64_t *seg_start, uint32_t blocks, uint64_t slots)
{
    if (threadIdx.x == 0) {
        seg_b0[0] = 0;
        seg_b0[1] = blocks;
        seg_start[0] = 0;
        seg_start[1] = slots;
    }
}

// counts[(segment, bin, workgroup)] = slots of the workgroup's range in bin.
__global__ __launch_bounds__(kThreads) void part_count_kernel(const int32_t *in, PartPlan pl, uint64_t n,
                                                              uint64_t *counts)
{
    __shared__ uint32_t hist[kMaxBins];
    if (blockIdx.x >= pl.seg_b0[pl.nseg]) return;
    const BlockRange r = block_range(pl, blockIdx.x);
    for (uint32_t i = threadIdx.x; i < pl.nbins; i += kThreads) hist[i] = 0;
    __syncthreads();
    // kPerThread reads in flight per thread; the next tile's are issued
    // before this tile's LDS updates
    uint32_t hot = 0;  // this lane's slots in bin 0
    // a tile inside the range loads from a uniform base with 32-bit offsets
    auto load_tile = [&](uint64_t t, int32_t *x) {
        if (t + kTile <= r.hi) {
            const int32_t *b = in + t;
#pragma unroll
            for (int j = 0; j < kPerThread; ++j) x[j] = b[j * kThreads + threadIdx.x];
        } else {
#pragma unroll
            for (int j = 0; j < kPerThread; ++j) {
                const uint64_t i = t + (uint64_t)j * kThreads + threadIdx.x;
                x[j] = i < r.hi ? in[i] : -1;
            }
        }
    };
    int32_t v[kPerThread];
    load_tile(r.lo, v);
    for (uint64_t t0 = r.lo; t0 < r.hi; t0 += kTile) {
        int32_t nv[kPerThread];
        load_tile(t0 + kTile, nv);
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
    for (uint32_t bin = threadIdx.x; bin < pl.nbins; bin += kThreads)
        counts[r.count_base + (uint64_t)bin * r.nb + r.rank] = hist[bin];
}

// Exclusive scan of counts[0, len) into offsets[0, len], offsets[len] = total:
// workgroup sums over chunks of `chunk` entries, one top-level scan, then
// each workgroup scans its chunk kScanChunk entries at a time.
template <typename C>
__global__ __launch_bounds__(kThreads) void scan_reduce_kernel(const C *counts, uint64_t len, uint64_t chunk,
                                                               uint64_t *partials)
{
    __shared__ uint64_t red[kThreads / 64];
    const uint64_t base = (uint64_t)blockIdx.x * chunk;
    const uint64_t end = base + chunk < len ? base + chunk : len;
    uint64_t s = 0;
    for (uint64_t i = base + threadIdx.x; i < end; i += kThreads) s += counts[i];
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

template <typename C>
__global__ __launch_bounds__(kThreads) void scan_apply_kernel(const C *counts, uint64_t len, uint64_t chunk,
                                                              const uint64_t *partials, uint64_t *offsets)
{
    __shared__ uint64_t red[kThreads / 64];
    constexpr uint32_t kPer = kScanChunk / kThreads;  // consecutive entries per thread
    const uint64_t start = (uint64_t)blockIdx.x * chunk;
    const uint64_t end = start + chunk < len ? start + chunk : len;
    uint64_t carry = partials[blockIdx.x];
    for (uint64_t step = start; step < end; step += kScanChunk) {
        const uint64_t base = step + (uint64_t)threadIdx.x * kPer;
        uint64_t v[kPer];
        uint64_t s = 0;
#pragma unroll
        for (uint32_t j = 0; j < kPer; ++j) {
            v[j] = base + j < end ? (uint64_t)counts[base + j] : 0;
            s += v[j];
        }
        uint64_t total = 0;
        uint64_t run = block_exclusive_scan<kThreads>(s, &total, red) + carry;
#pragma unroll
        for (uint32_t j = 0; j < kPer; ++j) {
            if (base + j < end) offsets[base + j] = run;
            run += v[j];
        }
        carry += total;
    }
}

uint32_t grid_of(uint64_t work, uint32_t block, uint32_t cap)
{
    uint64_t g = (work + block - 1) / block;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(g, cap));
}

// offsets[0, len] = exclusive scan of counts[0, len) with offsets[len] the
// total; partials holds kScanParts entries.
template <typename C>
hipError_t scan_counts(const C *counts, uint64_t len, uint64_t *partials, uint64_t *offsets, hipStream_t stream)
{
    uint64_t chunk = kScanChunk;
    while ((len + chunk - 1) / chunk > kScanParts) chunk *= 2;
    const uint32_t parts = (uint32_t)std::max<uint64_t>(1, (len + chunk - 1) / chunk);
    hipLaunchKernelGGL(scan_reduce_kernel<C>, dim3(parts), dim3(kThreads), 0, stream, counts, len, chunk, partials);
    hipLaunchKernelGGL(scan_top_kernel, dim3(1), dim3(kScanParts), 0, stream, partials, parts, offsets + len);
    hipLaunchKernelGGL(scan_apply_kernel<C>, dim3(parts), dim3(kThreads), 0, stream, counts, len, chunk, partials,
                       offsets);
    return hipGetLastError();
}

hipError_t scan(const uint64_t *counts, uint64_t len, const DegreeWork &w, hipStream_t stream)
{
    return scan_counts(counts, len, w.partials, w.offsets, stream);
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// ---- degree pass --------------------------------------------------------------

struct DegGeometry {
    uint64_t first = 0, slots = 0;  // generated slots [first, first + slots) of G
    uint32_t levels = 0;            // partition levels: 0, 1 or 2
    uint32_t buckets = 0;           // level-1 bins
    uint32_t shift = 0;             // level-1 bin of v: v >> shift
    uint32_t blocks = 1;            // partition workgroups W
    uint32_t leaves = 1;
    uint64_t len1 = 0, len2 = 0;    // entries of the level-1 counts (buckets x W) and of the matrix (leaves x W)
};

DegGeometry deg_geometry(uint64_t n, uint32_t d, uint32_t num_cus)
{
    DegGeometry g;
    g.first = (uint64_t)(d + 1) * d;
    g.slots = n > (uint64_t)d + 1 ? (n - d - 1) * (uint64_t)d : 0;
    // kPartPerCU partition workgroups per CU; the count kernel holds one per CU
    // (128 KB of LDS), the scatters two (88-96 VGPRs), so both run in rounds
    const uint64_t b = std::min<uint64_t>((uint64_t)num_cus * kPartPerCU, kMaxBlocks);
    g.blocks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(b, (g.slots + 65535) / 65536));
    const uint64_t leaves = (n + kLeafNodes - 1) >> kLeafBits;
    if (leaves <= 1) {
        g.levels = 0;
        g.leaves = 1;
    } else if (leaves <= kMaxBuckets) {
        g.levels = 1;
        g.buckets = (uint32_t)leaves;
        g.shift = kLeafBits;
        g.leaves = g.buckets;
        g.len1 = (uint64_t)g.buckets * g.blocks;
    } else {
        g.levels = 2;
        g.buckets = (uint32_t)((n + (1ull << kBucketShift) - 1) >> kBucketShift);
        g.shift = kBucketShift;
        g.leaves = g.buckets * kSubBins;
        g.len1 = (uint64_t)g.buckets * g.blocks;
        g.len2 = (uint64_t)g.leaves * g.blocks;
    }
    return g;
}

// Slots [lo, hi) of the input that partition workgroup w of W handles; the
// count and the level-1 scatter share them.
__device__ __forceinline__ void deg_range(uint64_t slots, uint32_t W, uint32_t w, uint64_t *lo, uint64_t *hi)
{
    const uint64_t per = (slots + W - 1) / W;
    const uint64_t a = (uint64_t)w * per;
    *lo = a < slots ? a : slots;
    *hi = a + per < slots ? a + per : slots;
}

// Per workgroup w: c1[b * W + w] = its valid slots in level-1 bucket b and,
// with MAT, cnt2[(b * 256 + s) * W + w] = those in leaf s of bucket b, for
// the buckets [b_lo, b_lo + nb), nb <= 128. The hub nodes v < hub_n (the
// smallest ids, which a preferential-attachment graph's slots name most: a
// tenth of them below 8192 at n = 2^30) are counted here, exactly, in LDS and
// added to hub[v], and are left out of the buckets, so the scatters and the
// leaf counts never see them. Node 0 and cell 0 are counted per lane instead
// of contending for one LDS word.
template <bool MAT>
__global__ __launch_bounds__(kCountThreads) void deg_count_kernel(const int32_t *in, uint64_t slots, uint64_t n,
                                                                  uint32_t shift, uint32_t buckets, uint32_t b_lo,
                                                                  uint32_t nb, uint32_t hub_n, uint32_t *c1,
                                                                  uint32_t *cnt2, uint32_t *hub, int32_t *status)
{
    constexpr uint32_t kCells = MAT ? kMatBuckets * kSubBins : kMaxBuckets;
    constexpr int kT = kCountThreads * kCountPer;
    __shared__ uint32_t cell[kCells];
    __shared__ uint32_t hubc[kHubNodes];
    const uint32_t W = gridDim.x, w = blockIdx.x;
    uint64_t lo, hi;
    deg_range(slots, W, w, &lo, &hi);
    const uint32_t ncell = MAT ? nb * kSubBins : buckets;
    for (uint32_t i = threadIdx.x; i < ncell; i += kCountThreads) cell[i] = 0;
    for (uint32_t i = threadIdx.x; i < hub_n; i += kCountThreads) hubc[i] = 0;
    __syncthreads();
    uint32_t hot = 0, hot_hub = 0;
    // kCountPer reads in flight per thread; the next tile's are issued
    // before this tile's LDS updates. A tile inside the range loads from a
    // uniform base with 32-bit offsets.
    auto load_tile = [&](uint64_t t, int32_t *x) {
        if (t + kT <= hi) {
            const int32_t *b = in + t;
#pragma unroll
            for (int j = 0; j < kCountPer; ++j) x[j] = b[j * kCountThreads + threadIdx.x];
        } else {
#pragma unroll
            for (int j = 0; j < kCountPer; ++j) {
                const uint64_t i = t + (uint64_t)j * kCountThreads + threadIdx.x;
                x[j] = i < hi ? in[i] : -1;
            }
        }
    };
    int32_t v[kCountPer];
    load_tile(lo, v);
    for (uint64_t t0 = lo; t0 < hi; t0 += kT) {
        int32_t nv[kCountPer];
        load_tile(t0 + kT, nv);
#pragma unroll
        for (int j = 0; j < kCountPer; ++j) {
            if (!valid(v[j], n)) continue;
            const uint32_t u = (uint32_t)v[j];
            if (u < hub_n) {
                if (u == 0) ++hot_hub;
                else atomicAdd(&hubc[u], 1u);
                continue;
            }
            uint32_t c = u >> shift;
            if (MAT) {
                const uint32_t b = c - b_lo;  // wraps below b_lo
                if (b >= nb) continue;
                c = b * kSubBins + ((u >> kLeafBits) & (kSubBins - 1));
            }
            if (c == 0) ++hot;
            else atomicAdd(&cell[c], 1u);
        }
#pragma unroll
        for (int j = 0; j < kCountPer; ++j) v[j] = nv[j];
    }
    hot = wave_sum(hot);
    if ((threadIdx.x & 63u) == 0 && hot) atomicAdd(&cell[0], hot);
    hot_hub = wave_sum(hot_hub);
    if ((threadIdx.x & 63u) == 0 && hot_hub) atomicAdd(&hubc[0], hot_hub);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < hub_n; i += kCountThreads)
        if (hubc[i]) check_u32_add(atomicAdd(&hub[i], hubc[i]), hubc[i], status);
    if (MAT) {
        for (uint32_t i = threadIdx.x; i < ncell; i += kCountThreads)
            cnt2[((uint64_t)b_lo * kSubBins + i) * W + w] = cell[i];
        // bucket sums; thread b starts in its own bank
        for (uint32_t b = threadIdx.x; b < nb; b += kCountThreads) {
            uint32_t s = 0;
            for (uint32_t k = 0; k < kSubBins; ++k) s += cell[b * kSubBins + ((k + b) & (kSubBins - 1))];
            c1[(uint64_t)(b_lo + b) * W + w] = s;
        }
    } else {
        for (uint32_t b = threadIdx.x; b < buckets; b += kCountThreads) c1[(uint64_t)b * W + w] = cell[b];
    }
}

// Moves the valid keys of in[lo, hi) to their bins' runs of the output, one
// 12288-slot tile at a time: the tile is grouped by bin ((v >> shift) & mask,
// at most 256 bins) in LDS first, so each bin's run leaves the workgroup as
// contiguous stores (at 128-256 bins a run averages 48-96 slots). At
// n = 2^30, d = 8 the two scatters took 18.3 + 17.3 ms with 16 slots per
// thread (8192-slot tiles), 17.3 + 15.8 with 24, 17.4 + 15.4 with 28 (125
// VGPRs, at the edge of four wavefronts per SIMD) and 26.9 + 22.5 with 32
// (143 VGPRs). cur[] holds
// each bin's next output position and thist[] must be zero on entry; OUT16
// stores a key's low 15 bits (the last level), otherwise the whole key. Bin
// 0 is ranked with one LDS atomic per wavefront. Ends with a barrier.
template <bool OUT16>
__device__ __forceinline__ void scatter_run(const int32_t *in, uint64_t lo, uint64_t hi, uint64_t n, uint32_t key_lo,
                                            uint32_t shift, uint32_t mask, uint32_t nbins, uint64_t *cur,
                                            uint32_t *thist, uint32_t *tstart, int32_t *tile, uint64_t *red,
                                            int32_t *out32, uint16_t *out16)
{
    // keys in [key_lo, n) move; the rest (out of range, or hub nodes the count pass took) stay
    auto keep = [&](int32_t x) { return valid(x, n) && (uint32_t)x >= key_lo; };
    // a tile inside the range loads from a uniform base with 32-bit offsets
    auto load_tile = [&](uint64_t t, int32_t *x) {
        if (t + kPartTile <= hi) {
            const int32_t *p = in + t;
#pragma unroll
            for (int j = 0; j < kPartPer; ++j) x[j] = p[j * kPartThreads + threadIdx.x];
        } else {
#pragma unroll
            for (int j = 0; j < kPartPer; ++j) {
                const uint64_t i = t + (uint64_t)j * kPartThreads + threadIdx.x;
                x[j] = i < hi ? in[i] : -1;
            }
        }
    };
    int32_t v[kPartPer];
    load_tile(lo, v);
    const uint32_t b = threadIdx.x;  // this thread's bin in the per-tile scan
    for (uint64_t t0 = lo; t0 < hi; t0 += kPartTile) {
        uint32_t rank[kPartPer];
#pragma unroll
        for (int j = 0; j < kPartPer; ++j) {
            const bool ok = keep(v[j]);
            const uint32_t bin = ((uint32_t)v[j] >> shift) & mask;
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
        const uint32_t c = b < nbins ? thist[b] : 0u;
        uint64_t tot = 0;
        const uint32_t ex = (uint32_t)block_exclusive_scan<kPartThreads>(c, &tot, red);
        if (b < nbins) tstart[b] = ex;
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kPartPer; ++j)
            if (keep(v[j])) tile[tstart[((uint32_t)v[j] >> shift) & mask] + rank[j]] = v[j];
        // the next tile's reads are in flight while this one is stored
        load_tile(t0 + kPartTile, v);
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < (uint32_t)tot; i += kPartThreads) {
            const int32_t x = tile[i];
            const uint32_t bin = ((uint32_t)x >> shift) & mask;
            const uint64_t o = cur[bin] + (i - tstart[bin]);
            if (OUT16) out16[o] = (uint16_t)((uint32_t)x & (kLeafNodes - 1));
            else out32[o] = x;
        }
        __syncthreads();
        if (b < nbins) {
            cur[b] += thist[b];
            thist[b] = 0;
        }
        __syncthreads();
    }
}

// Level 1: workgroup w's slots (the count kernel's range) to the buckets,
// its run of bucket b starting at offs1[b * W + w].
template <bool OUT16>
__global__ __launch_bounds__(kPartThreads) void deg_scatter1_kernel(const int32_t *in, uint64_t slots, uint64_t n,
                                                                   uint32_t hub_n, uint32_t shift, uint32_t buckets,
                                                                   const uint64_t *offs1, int32_t *out32,
                                                                   uint16_t *out16)
{
    __shared__ uint64_t cur[kMaxBuckets];
    __shared__ uint32_t thist[kMaxBuckets], tstart[kMaxBuckets];
    __shared__ int32_t tile[kPartTile];
    __shared__ uint64_t red[kPartThreads / 64];
    const uint32_t W = gridDim.x, w = blockIdx.x;
    uint64_t lo, hi;
    deg_range(slots, W, w, &lo, &hi);
    for (uint32_t b = threadIdx.x; b < buckets; b += kPartThreads) {
        cur[b] = offs1[(uint64_t)b * W + w];
        thist[b] = 0;
    }
    __syncthreads();
    scatter_run<OUT16>(in, lo, hi, n, hub_n, shift, kMaxBuckets - 1, buckets, cur, thist, tstart, tile, red, out32,
                       out16);
}

// Level 2: workgroup j moves the run that level-1 workgroup j left in each
// bucket to the bucket's 256 leaves; the scanned matrix gives its run of
// leaf s of bucket b at offs2[(b * 256 + s) * W + j].
__global__ __launch_bounds__(kPartThreads) void deg_scatter2_kernel(const int32_t *part, uint64_t n, uint32_t buckets,
                                                                   const uint64_t *offs1, const uint64_t *offs2,
                                                                   uint16_t *out16)
{
    __shared__ uint64_t cur[kSubBins];
    __shared__ uint32_t thist[kSubBins], tstart[kSubBins];
    __shared__ int32_t tile[kPartTile];
    __shared__ uint64_t red[kPartThreads / 64];
    const uint32_t W = gridDim.x, j = blockIdx.x;
    for (uint32_t b = 0; b < buckets; ++b) {
        const uint64_t lo = offs1[(uint64_t)b * W + j], hi = offs1[(uint64_t)b * W + j + 1];
        if (lo == hi) continue;
        if (threadIdx.x < kSubBins) {
            cur[threadIdx.x] = offs2[((uint64_t)b * kSubBins + threadIdx.x) * W + j];
            thist[threadIdx.x] = 0;
        }
        __syncthreads();
        scatter_run<true>(part, lo, hi, n, 0u, kLeafBits, kSubBins - 1, kSubBins, cur, thist, tstart, tile, red,
                          nullptr, out16);
    }
}

__global__ void deg_whole_kernel(uint64_t *offs, uint64_t slots)
{
    if (threadIdx.x == 0) {
        offs[0] = 0;
        offs[1] = slots;
    }
}

__device__ __forceinline__ uint64_t pieces_of(uint64_t size) { return size > kPiece ? (size + kPiece - 1) / kPiece : 1; }

// piece_base = exclusive scan over the leaves of their pieces, leaf l holding
// the keys [offs[l * stride], offs[(l + 1) * stride]); one workgroup.
__global__ __launch_bounds__(kPlanThreads) void deg_leaf_plan_kernel(const uint64_t *offs, uint32_t stride,
                                                                    uint32_t leaves, uint64_t *piece_base)
{
    __shared__ uint64_t red[kPlanThreads / 64];
    const uint32_t per = (leaves + kPlanThreads - 1) / kPlanThreads;
    const uint32_t lo = std::min<uint32_t>(threadIdx.x * per, leaves), hi = std::min<uint32_t>(lo + per, leaves);
    auto size = [&](uint32_t l) { return offs[(uint64_t)(l + 1) * stride] - offs[(uint64_t)l * stride]; };
    uint64_t p = 0;
    for (uint32_t l = lo; l < hi; ++l) p += pieces_of(size(l));
    uint64_t tot = 0;
    uint64_t run = block_exclusive_scan<kPlanThreads>(p, &tot, red);
    for (uint32_t l = lo; l < hi; ++l) {
        piece_base[l] = run;
        run += pieces_of(size(l));
    }
    if (threadIdx.x == 0) piece_base[leaves] = tot;
}

// deg = d over the leaves counted in several pieces, which add their counts
// with atomics.
__global__ void deg_hub_init_kernel(const uint64_t *piece_base, uint32_t leaves, uint64_t n, uint32_t d, uint32_t *deg)
{
    for (uint32_t l = blockIdx.x; l < leaves; l += gridDim.x) {
        if (piece_base[l + 1] - piece_base[l] <= 1) continue;
        const uint64_t base = (uint64_t)l << kLeafBits;
        for (uint32_t i = threadIdx.x; i < kLeafNodes; i += blockDim.x)
            if (base + i < n) deg[base + i] = d;
    }
}

// The hub nodes' counts from the count pass, added after the leaf counts
// (which never saw their slots).
__global__ void deg_hub_add_kernel(const uint32_t *hub, uint32_t hub_n, uint32_t *deg, int32_t *status)
{
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < hub_n; v += gridDim.x * blockDim.x) {
        check_u32_add(deg[v], hub[v], status);
        deg[v] += hub[v];
    }
}

__global__ void zero_u32_kernel(uint32_t *p, uint32_t len)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < len; i += gridDim.x * blockDim.x) p[i] = 0;
}

// deg of every node of a leaf from LDS counters: a one-piece leaf stores
// d + count, a split leaf adds its piece's counts to deg. K16: the leaf keys
// are 16-bit (read 8 per 16-byte load); otherwise the graph's own int32 slots
// (a graph of at most 2^15 nodes, one leaf), which are checked. The leaf's
// first node, a hub in leaf 0, is counted per lane.
template <bool K16>
__global__ __launch_bounds__(kLeafThreads) void deg_leaf_kernel(const void *keys, const uint64_t *offs,
                                                               uint32_t stride, uint32_t leaves,
                                                               const uint64_t *piece_base, uint64_t n, uint32_t d,
                                                               uint32_t *deg, int32_t *status)
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
        const uint64_t lend = offs[(uint64_t)(lo + 1) * stride];
        const uint64_t first = offs[(uint64_t)lo * stride] + (q - piece_base[lo]) * kPiece;
        const uint64_t end = first + kPiece < lend ? first + kPiece : lend;
        for (uint32_t i = threadIdx.x; i < kLeafNodes; i += kLeafThreads) cnt[i] = 0;
        __syncthreads();
        uint32_t hot = 0;
        auto count = [&](uint32_t x) {
            if (x == 0) ++hot;
            else if (x < kLeafNodes) atomicAdd(&cnt[x], 1u);
        };
        if (K16) {
            const uint16_t *k16 = static_cast<const uint16_t *>(keys);
            const uint4 *k4 = static_cast<const uint4 *>(keys);  // 256-byte aligned buffer
            const uint64_t a0 = std::min<uint64_t>(end, (first + 7) & ~7ull);  // head up to 8-key alignment
            const uint64_t a1 = std::max<uint64_t>(a0, end & ~7ull);           // 16-byte body [a0, a1)
            for (uint64_t i = first + threadIdx.x; i < a0; i += kLeafThreads) count(k16[i]);
            const uint64_t g1 = a1 / 8;
            for (uint64_t g = a0 / 8 + threadIdx.x; g < g1; g += kLeafUnroll * kLeafThreads) {
                uint4 x[kLeafUnroll];
#pragma unroll
                for (int u = 0; u < kLeafUnroll; ++u) {
                    const uint64_t gg = g + (uint64_t)u * kLeafThreads;
                    x[u] = gg < g1 ? k4[gg] : make_uint4(~0u, ~0u, ~0u, ~0u);  // 0xffff: no key
                }
#pragma unroll
                for (int u = 0; u < kLeafUnroll; ++u) {
                    const uint32_t wd[4] = {x[u].x, x[u].y, x[u].z, x[u].w};
#pragma unroll
                    for (int h = 0; h < 4; ++h) {
                        count(wd[h] & 0xffffu);
                        count(wd[h] >> 16);
                    }
                }
            }
            for (uint64_t i = a1 + threadIdx.x; i < end; i += kLeafThreads) count(k16[i]);
        } else {
            const int32_t *k32 = static_cast<const int32_t *>(keys);
            for (uint64_t i = first + threadIdx.x; i < end; i += kLeafThreads) {
                const int32_t v = k32[i];
                if (valid(v, n)) count((uint32_t)v);
            }
        }
        hot = wave_sum(hot);
        if ((threadIdx.x & 63u) == 0 && hot) atomicAdd(&cnt[0], hot);
        __syncthreads();
        const uint64_t base = (uint64_t)lo << kLeafBits;
        for (uint32_t i = threadIdx.x; i < kLeafNodes; i += kLeafThreads) {
            const uint64_t v = base + i;
            if (v < n) {
                if (np == 1) deg[v] = d + cnt[i];  // at most kPiece slots: no overflow
                else if (cnt[i]) check_u32_add(atomicAdd(&deg[v], cnt[i]), cnt[i], status);
            }
        }
        __syncthreads();
    }
}

// ---- in-edge CSR: a stable LSD radix sort of the slots by target ------------
// Pass p orders the (target, source) pairs by its digit of the target (the
// target's bits split over ceil(bits(n-1) / 8) passes in digits of at most 8
// bits), keeping the order of equal digits; after the last pass the sources
// of each target are in slot order, which is ascending source order.
// Ranks are exact without LDS atomics: a wavefront owns 1024 consecutive
// slots of a tile (8192 slots for the 8 wavefronts of a workgroup) and walks
// them 64 at a time; lanes with equal digits are matched with 8 ballots and
// counted from a wavefront-private running count, and the wavefronts' counts
// are combined per digit. A digit's run in a tile averages 32 pairs at 8
// bits and 64 at 7, so most of its stores fill whole lines.
constexpr uint32_t kRadixBits = 8;
constexpr uint32_t kRadixBins = 1u << kRadixBits;
constexpr uint32_t kLsdThreads = 512;
constexpr uint32_t kWaveShare = 1024;             // slots per wavefront per tile
constexpr uint32_t kSteps = kWaveShare / 64;      // 64-slot steps per wavefront

// Lanes among `act` whose digit equals this lane's; whole wavefront.
// Per bit: a sign-extended bit field s (0 or all ones), its ballot m, and
// each half of the mask ANDed with ~(m ^ s) (lanes whose bit equals this
// lane's): six VALU instructions, where a per-lane select between m and ~m
// of 64-bit values took ten.
__device__ __forceinline__ uint64_t match_digit(uint32_t digit, bool act)
{
    const uint64_t a = __ballot(act);
    uint32_t lo = (uint32_t)a, hi = (uint32_t)(a >> 32);
#pragma unroll
    for (uint32_t b = 0; b < kRadixBits; ++b) {
        const uint32_t s = (uint32_t)((int32_t)(digit << (31 - b)) >> 31);
        const uint64_t m = __ballot(s != 0u);
        lo &= ~((uint32_t)m ^ s);
        hi &= ~((uint32_t)(m >> 32) ^ s);
    }
    return ((uint64_t)hi << 32) | lo;
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

// One stable pass. FIRST: the keys are the targets array itself and the
// sources the slots' rows i / d (vin unused); otherwise both are read. The
// keys are written for the next pass (the last pass's, for in_ptr). A wavefront whose
// 1024 slots lie inside the range loads them from a wavefront-uniform base
// with 32-bit offsets and no bound checks; the first pass steps each lane's
// row i / d by 64 slots at a time (qstep = 64 / d, rstep = 64 % d).
template <uint32_t TH, bool FIRST>
__global__ __launch_bounds__(TH) __attribute__((amdgpu_waves_per_eu(4))) void lsd_scatter_kernel(const int32_t *kin, const int32_t *vin, int32_t *kout,
                                                         int32_t *vout, PartPlan pl, uint64_t n, uint32_t d,
                                                         double inv_d, uint32_t qstep, uint32_t rstep,
                                                         const uint64_t *offsets)
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
    const uint32_t lane = lane_id();
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const bool digit_thread = threadIdx.x < kRadixBins;
    if (digit_thread) cur[threadIdx.x] = offsets[r.count_base + (uint64_t)threadIdx.x * r.nb + r.rank];
    __syncthreads();
    for (uint64_t t0 = r.lo; t0 < r.hi; t0 += kT) {
        for (uint32_t b = lane; b < kRadixBins; b += 64) wrun[wave][b] = 0;
        int32_t k[kSteps], v[kSteps];
        uint32_t rk[kSteps];
        const uint64_t w0 = t0 + (uint64_t)wave * kWaveShare;  // this wavefront's first slot
        if (w0 + kWaveShare <= r.hi) {
            const int32_t *kb = kin + w0;
            const int32_t *vb = vin + w0;
#pragma unroll
            for (uint32_t j = 0; j < kSteps; ++j) {
                k[j] = kb[j * 64 + lane];
                if (!FIRST) v[j] = vb[j * 64 + lane];
            }
        } else {
#pragma unroll
            for (uint32_t j = 0; j < kSteps; ++j) {
                const uint64_t i = w0 + j * 64 + lane;
                k[j] = i < r.hi ? kin[i] : -1;
                if (!FIRST) v[j] = i < r.hi ? vin[i] : 0;
            }
        }
        if (FIRST) {
            uint32_t q = (uint32_t)div_small(w0 + lane, d, inv_d);
            uint32_t rem = (uint32_t)(w0 + lane - (uint64_t)q * d);
#pragma unroll
            for (uint32_t j = 0; j < kSteps; ++j) {
                v[j] = (int32_t)q;
                q += qstep;
                rem += rstep;
                if (rem >= d) {
                    rem -= d;
                    ++q;
                }
            }
        }
#pragma unroll
        for (uint32_t j = 0; j < kSteps; ++j) {
            const bool act = valid(k[j], n);
            const uint32_t dig = ((uint32_t)k[j] >> pl.shift) & pl.mask;
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
                const uint32_t dig = ((uint32_t)k[j] >> pl.shift) & pl.mask;
                const uint32_t pos = tstart[dig] + wrun[wave][dig] + rk[j];
                tk[pos] = k[j];
                tv[pos] = v[j];
            }
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < (uint32_t)total; i += TH) {
            const int32_t kk = tk[i];
            const uint32_t dig = ((uint32_t)kk >> pl.shift) & pl.mask;
            const uint64_t o = cur[dig] + (i - tstart[dig]);
            kout[o] = kk;
            vout[o] = tv[i];
        }
        __syncthreads();
        if (digit_thread) cur[threadIdx.x] += tot;
        __syncthreads();
    }
}

// in_ptr[v] = first position of target v in the sorted keys: position i
// (0 <= i <= total, the key at `total` taken as n) writes in_ptr[v] = i for
// every v in (key[i-1], key[i]], so in_ptr[n] = total. A wavefront walks
// spans of 16 x 64 consecutive positions, lane l taking position j*64 + l of
// step j: the loads are coalesced, the key before a position comes from the
// lane before (or the step before's last lane, or one load for the span's
// first), and consecutive lanes write consecutive in_ptr entries. (Four
// consecutive positions per thread with an extra load of the key before them
// took 14.5 ms at n = 2^30, d = 8.) A lane whose range (key[i-1], key[i]]
// holds more than kPtrLane nodes (nodes no slot names: the tail of a graph
// whose last nodes have no in-edges, or any targets array with few distinct
// values) leaves it to the whole wavefront, which writes 64 entries a store.
constexpr uint32_t kPtrSteps = 16;  // 9.2 ms at n = 2^30, d = 8 (8 steps: 9.4, 32: 10.1)
constexpr int64_t kPtrLane = 32;
__global__ __launch_bounds__(256) void csr_ptr_kernel(const int32_t *keys, const uint64_t *total_p, uint64_t n,
                                                      uint64_t cap, int64_t *in_ptr)
{
    const uint64_t total = *total_p < cap ? *total_p : cap;
    const uint64_t positions = total + 1;
    const uint32_t lane = lane_id();
    const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / 64);
    const uint64_t wave = (uint64_t)blockIdx.x * (blockDim.x / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    constexpr uint64_t kSpan = 64ull * kPtrSteps;
    for (uint64_t s = wave * kSpan; s < positions; s += waves * kSpan) {
        int32_t k[kPtrSteps];
#pragma unroll
        for (uint32_t j = 0; j < kPtrSteps; ++j) {
            const uint64_t i = s + j * 64 + lane;
            k[j] = i < total ? keys[i] : (int32_t)n;  // n at position total (and unused beyond)
        }
        int32_t before = s == 0 ? -1 : keys[s - 1];  // the key before the span (one address for all lanes)
#pragma unroll
        for (uint32_t j = 0; j < kPtrSteps; ++j) {
            const uint64_t i = s + j * 64 + lane;
            const int32_t up = __shfl_up(k[j], 1, kWave);
            const int32_t prev = lane == 0 ? before : up;
            const int64_t lo = (int64_t)prev + 1, hi = k[j];  // in_ptr[lo..hi] = i
            const bool wide = i <= total && hi - lo >= kPtrLane;
            if (i <= total && !wide)
                for (int64_t v = lo; v <= hi; ++v) in_ptr[v] = (int64_t)i;
            for (uint64_t m = __ballot(wide); m != 0; m &= m - 1) {
                const int l = (int)__builtin_ctzll(m);
                // a wide lane's lo and hi are node ids <= n <= 2^31 - 1
                const int64_t wlo = __builtin_amdgcn_readlane((int)lo, l);
                const int64_t whi = __builtin_amdgcn_readlane((int)hi, l);
                const int64_t wi = (int64_t)(s + j * 64) + l;
                for (int64_t v = wlo + lane; v <= whi; v += 64) in_ptr[v] = wi;
            }
            before = __shfl(k[j], 63, kWave);
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
    uint32_t bits = 1, passes = 1, blocks = 1;
    uint64_t counts = 0;
};

// Digit widths: the key's bits split as evenly as the passes allow, the
// spare bits of an uneven split going to the high digits. At n = 2^30
// (30 bits, 4 passes) widths 7,7,8,8 build the in-edge CSR in 185.6 ms;
// 6,8,8,8 in 187.7, 7,8,8,7 in 187.5, 8,8,7,7 in 191.5 and 8,8,8,6 in
// 195.9: the low digits are uniform, and 128 bins leave runs of 64 pairs per
// tile, whole lines; the high digits are skewed towards the small ids, which
// already lengthens the runs of the bins that hold most of the pairs.
void digit_of_pass(const CsrGeometry &g, uint32_t pass, uint32_t *shift, uint32_t *width)
{
    const uint32_t base = g.bits / g.passes, extra = g.bits % g.passes;
    auto w = [&](uint32_t i) -> uint32_t { return base + (i >= g.passes - extra ? 1u : 0u); };
    uint32_t sh = 0;
    for (uint32_t i = 0; i < pass; ++i) sh += w(i);
    *shift = sh;
    *width = w(pass);
}

CsrGeometry csr_geometry(uint64_t n, uint32_t d, uint32_t num_cus)
{
    CsrGeometry g;
    g.slots = n * (uint64_t)d;
    uint32_t bits = 1;
    while (bits < 63 && (1ull << bits) < n) ++bits;  // targets are < n <= 2^bits
    g.bits = bits;
    g.passes = (bits + kRadixBits - 1) / kRadixBits;
    // four workgroups per CU, two resident (n = 2^30, d = 8: two per CU 185 ms,
    // eight a median 177 against 180, inside the spread between processes)
    uint64_t b = std::min<uint64_t>((uint64_t)num_cus * 4, kMaxBlocks);
    g.blocks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(b, (g.slots + 65535) / 65536));
    g.counts = (uint64_t)kRadixBins * g.blocks + 1;
    return g;
}

}  // namespace

size_t degree_work_bytes(uint64_t n, uint32_t d, uint32_t num_cus)
{
    const DegGeometry g = deg_geometry(n, d, num_cus);
    size_t b = 0;
    if (g.levels == 2) b += align256(g.slots * 4);
    if (g.levels >= 1) b += align256(g.slots * 2);
    b += align256(g.len1 * 4) + align256((g.len1 + 1) * 8);
    b += align256(g.len2 * 4) + align256((g.len2 + 1) * 8);
    b += align256(kScanParts * 8) + align256(2 * 8) + align256(((size_t)g.leaves + 1) * 8);
    b += align256(kHubNodes * 4);
    return b;
}

void degree_work_layout(void *base, uint64_t n, uint32_t d, uint32_t num_cus, DegreeWork *w)
{
    const DegGeometry g = deg_geometry(n, d, num_cus);
    char *p = (char *)base;
    auto take = [&](size_t bytes) {
        char *q = p;
        p += align256(bytes);
        return (void *)q;
    };
    *w = DegreeWork{};
    // the 16-bit keys first: the level-1 scatter's writes into part_a took
    // 17.6 ms at the start of the work area and 15.9-16.4 ms after part_b
    // (n = 2^30, d = 8)
    if (g.levels >= 1) w->part_b = (uint16_t *)take(g.slots * 2);
    if (g.levels == 2) w->part_a = (int32_t *)take(g.slots * 4);
    w->c1 = (uint32_t *)take(g.len1 * 4);
    w->offs1 = (uint64_t *)take((g.len1 + 1) * 8);
    w->c2 = (uint32_t *)take(g.len2 * 4);
    w->offs2 = (uint64_t *)take((g.len2 + 1) * 8);
    w->partials = (uint64_t *)take(kScanParts * 8);
    w->whole = (uint64_t *)take(2 * 8);
    w->piece_base = (uint64_t *)take(((size_t)g.leaves + 1) * 8);
    w->hub = (uint32_t *)take(kHubNodes * 4);
    w->num_cus = num_cus;
}

hipError_t launch_degree_partitioned(const int32_t *G, uint64_t n, uint32_t d, uint32_t *deg, const DegreeWork &w,
                                     hipStream_t stream)
{
    const DegGeometry g = deg_geometry(n, d, w.num_cus);
    if (g.slots == 0) {
        hipLaunchKernelGGL(fill_deg_kernel, dim3(grid_of(n, kThreads, 8192)), dim3(kThreads), 0, stream, deg, n, d);
        return hipGetLastError();
    }
    const int32_t *in = G + g.first;
    const uint32_t W = g.blocks;
    const uint64_t *offs = w.whole;
    uint32_t stride = 1;
    hipError_t e = hipSuccess;
    if (g.levels == 0) {
        hipLaunchKernelGGL(deg_whole_kernel, dim3(1), dim3(64), 0, stream, w.whole, g.slots);
    } else if (g.levels == 1) {
        hipLaunchKernelGGL(zero_u32_kernel, dim3(kHubNodes / kThreads), dim3(kThreads), 0, stream, w.hub, kHubNodes);
        hipLaunchKernelGGL(deg_count_kernel<false>, dim3(W), dim3(kCountThreads), 0, stream, in, g.slots, n, g.shift,
                           g.buckets, 0u, 0u, kHubNodes, w.c1, (uint32_t *)nullptr, w.hub, w.status);
        if ((e = scan_counts<uint32_t>(w.c1, g.len1, w.partials, w.offs1, stream)) != hipSuccess) return e;
        hipLaunchKernelGGL(deg_scatter1_kernel<true>, dim3(W), dim3(kPartThreads), 0, stream, in, g.slots, n,
                           kHubNodes, g.shift, g.buckets, (const uint64_t *)w.offs1, (int32_t *)nullptr, w.part_b);
        offs = w.offs1;
        stride = W;
    } else {
        hipLaunchKernelGGL(zero_u32_kernel, dim3(kHubNodes / kThreads), dim3(kThreads), 0, stream, w.hub, kHubNodes);
        for (uint32_t b0 = 0; b0 < g.buckets; b0 += kMatBuckets)  // the hub lies in bucket 0
            hipLaunchKernelGGL(deg_count_kernel<true>, dim3(W), dim3(kCountThreads), 0, stream, in, g.slots, n,
                               g.shift, g.buckets, b0, std::min<uint32_t>(kMatBuckets, g.buckets - b0),
                               b0 == 0 ? kHubNodes : 0u, w.c1, w.c2, w.hub, w.status);
        if ((e = scan_counts<uint32_t>(w.c1, g.len1, w.partials, w.offs1, stream)) != hipSuccess) return e;
        if ((e = scan_counts<uint32_t>(w.c2, g.len2, w.partials, w.offs2, stream)) != hipSuccess) return e;
        hipLaunchKernelGGL(deg_scatter1_kernel<false>, dim3(W), dim3(kPartThreads), 0, stream, in, g.slots, n,
                           kHubNodes, g.shift, g.buckets, (const uint64_t *)w.offs1, w.part_a, (uint16_t *)nullptr);
        hipLaunchKernelGGL(deg_scatter2_kernel, dim3(W), dim3(kPartThreads), 0, stream, (const int32_t *)w.part_a, n,
                           g.buckets, (const uint64_t *)w.offs1, (const uint64_t *)w.offs2, w.part_b);
        offs = w.offs2;
        stride = W;
    }
    hipLaunchKernelGGL(deg_leaf_plan_kernel, dim3(1), dim3(kPlanThreads), 0, stream, offs, stride, g.leaves,
                       w.piece_base);
    hipLaunchKernelGGL(deg_hub_init_kernel, dim3(std::min<uint32_t>(g.leaves, 1024)), dim3(kThreads), 0, stream,
                       (const uint64_t *)w.piece_base, g.leaves, n, d, deg);
    const uint32_t grid = std::max<uint32_t>(1, w.num_cus);
    if (g.levels == 0) {
        hipLaunchKernelGGL(deg_leaf_kernel<false>, dim3(grid), dim3(kLeafThreads), 0, stream, (const void *)in, offs,
                           stride, g.leaves, (const uint64_t *)w.piece_base, n, d, deg, w.status);
    } else {
        hipLaunchKernelGGL(deg_leaf_kernel<true>, dim3(grid), dim3(kLeafThreads), 0, stream, (const void *)w.part_b,
                           offs, stride, g.leaves, (const uint64_t *)w.piece_base, n, d, deg, w.status);
        hipLaunchKernelGGL(deg_hub_add_kernel, dim3(kHubNodes / kThreads), dim3(kThreads), 0, stream,
                           (const uint32_t *)w.hub, kHubNodes, deg, w.status);  // n > 2^15 > kHubNodes here
    }
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
    hipLaunchKernelGGL(plan_whole_kernel, dim3(1), dim3(64), 0, stream, seg_b0, seg_start, g.blocks, g.slots);
    const int32_t *kin = G, *vin = nullptr;
    for (uint32_t pass = 0; pass < g.passes; ++pass) {
        uint32_t shift = 0, width = kRadixBits;
        digit_of_pass(g, pass, &shift, &width);
        const PartPlan pl{seg_b0, seg_start, 1u, kRadixBins, shift, (1u << width) - 1};
        const bool last = pass + 1 == g.passes;
        // Even passes (the first, and the 8-bit third at n = 2^30) write the
        // second pair of buffers. Measured over ten alternating processes at
        // n = 2^30, d = 8: 179 ms median against 187 the other way round (equal
        // in the runs where both took 171); writes into the first 64 GB of the
        // work area ran slower, for a reason not identified.
        int32_t *kout = (pass & 1) ? ka : kb;
        int32_t *vout = last ? in_src : ((pass & 1) ? va : vb);
        hipLaunchKernelGGL(part_count_kernel, dim3(g.blocks), dim3(kThreads), 0, stream, kin, pl, n, w.counts);
        if ((e = scan(w.counts, len, w, stream)) != hipSuccess) return e;
        const uint32_t qstep = 64 / d, rstep = 64 % d;
        if (pass == 0)
            hipLaunchKernelGGL((lsd_scatter_kernel<kLsdThreads, true>), dim3(g.blocks), dim3(kLsdThreads), 0,
                               stream, kin, vin, kout, vout, pl, n, d, inv_d, qstep, rstep,
                               (const uint64_t *)w.offsets);
        else
            hipLaunchKernelGGL((lsd_scatter_kernel<kLsdThreads, false>), dim3(g.blocks), dim3(kLsdThreads), 0,
                               stream, kin, vin, kout, vout, pl, n, d, inv_d, qstep, rstep,
                               (const uint64_t *)w.offsets);
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
