// ppa_kernels.hip -- gfx950 kernels of the copy-model generator.
//
// Reference kernels replaced here (src/PA/cuPPA/):
//   initial_graph_rrp (cuPPA_RRP.cu:51-76) and the per-stage
//   initialize_vertex_structures_rrp (cuPPA_RRP.cu:78-102)  -> fill_kernel
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
//     by a lane whose own dependencies are all final, so the grid always makes
//     progress, whatever the residency of workgroups.
//   * Slots are write-once: the fill kernel sets every generated slot to -1
//     before the launch, and the owner later stores the final value with an
//     agent-scope store. A plain (L2-cached) read therefore returns either -1
//     or the final value; only a -1 needs a coherent (agent-scope) re-read.
//   * On claiming a node, a lane draws attempt 0 of all d slots at once and
//     issues their d copy reads together (d-way memory-level parallelism);
//     the duplicate check then walks the slots in order from registers. Only a
//     rejected attempt (duplicate) or an empty source slot costs another
//     round trip. The reference instead parks placeholders -(v*M+k) in a
//     per-thread LDS queue and spins (cuPPA_RRP_1.cu:97-185).
//   * The per-node duplicate set is the node's own d registers: d <= 16 means
//     an unrolled compare against at most 15 values, no tree walk.
#include "ppa_kernels.hpp"
#include "ppa_model.hpp"

namespace ppa {
namespace {

constexpr int kWave = 64;
constexpr uint32_t kMaxRetry = 1u << 24;
constexpr int kErrRetry = -2;
constexpr int kErrTimeout = -5;

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t uniform64(uint64_t v)
{
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint32_t rank_below(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ int32_t load_coherent(const int32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void store_coherent(int32_t *p, int32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

__global__ void reset_scratch_kernel(DeviceScratch *s, uint64_t first_node)
{
    if (threadIdx.x == 0) {
        s->claim = first_node;
        s->status = 0;
        s->max_degree = 0;
    }
    if (threadIdx.x < kStatCount) s->stats[threadIdx.x] = 0;
}

// Sets generated slots to -1, writes the seed clique rows and row_ptr[t] = t*d.
__global__ void fill_kernel(FillArgs a)
{
    const uint64_t slots = a.n * a.d;
    const uint64_t clique = (uint64_t)(a.d + 1) * a.d;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // Vector body: int4 groups; the host guarantees G is 16-byte aligned.
    const uint64_t groups = slots / 4;
    int4 *G4 = reinterpret_cast<int4 *>(a.G);
    for (uint64_t g = tid; g < groups; g += stride) {
        const uint64_t e = g * 4;
        int4 v = make_int4(kEmpty, kEmpty, kEmpty, kEmpty);
        if (e < clique) {
            int32_t w[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint64_t s = e + i;
                w[i] = s < clique ? clique_target(s / a.d, (uint32_t)(s % a.d)) : kEmpty;
            }
            v = make_int4(w[0], w[1], w[2], w[3]);
        }
        G4[g] = v;
    }
    for (uint64_t s = groups * 4 + tid; s < slots; s += stride)
        a.G[s] = s < clique ? clique_target(s / a.d, (uint32_t)(s % a.d)) : kEmpty;
    if (a.row_ptr)
        for (uint64_t t = tid; t <= a.n; t += stride) a.row_ptr[t] = (int64_t)(t * a.d);
}

template <int D>
__global__ __launch_bounds__(256) void generate_kernel(GenArgs a)
{
    const uint32_t lane = lane_id();
    int32_t *const G = a.G;
    const uint64_t n = a.n;
    const uint64_t T = a.T;
    const uint64_t start_clock = wall_clock64();

    // Wavefront-uniform claim window [wcur, wend).
    uint64_t wcur = 0, wend = 0;
    bool wdone = false;

    // Lane state.
    uint64_t t = 0;
    bool active = false, fresh = false, need_draw = false, have_src = false;
    uint32_t k = 0, r = 0;
    uint64_t src = 0;
    int32_t val[D];
#pragma unroll
    for (int i = 0; i < D; ++i) val[i] = kEmpty;

    uint32_t c_nodes = 0, c_draws = 0, c_retries = 0, c_pmiss = 0, c_polls = 0, c_stalls = 0,
             c_claims = 0, c_sleeps = 0;
    uint32_t iter = 0, backoff = 1;
    int fail = 0;

    while (true) {
        // ---- claim nodes for idle lanes (wavefront-uniform control flow) ----
        bool progressed = false;
        const uint64_t idle_mask = __ballot(!active);
        if (!wdone && idle_mask != 0) {
            if (wcur >= wend) {
                uint64_t base = 0;
                if (lane == 0) {
                    base = atomicAdd(&a.scratch->claim, (unsigned long long)a.chunk);
                    ++c_claims;
                }
                base = uniform64(base);
                if (base >= n) {
                    wdone = true;
                } else {
                    wcur = base;
                    wend = base + a.chunk < n ? base + a.chunk : n;
                }
            }
            if (!wdone) {
                const uint64_t avail = wend - wcur;
                const uint32_t rank = rank_below(idle_mask);
                if (!active && rank < avail) {
                    t = wcur + rank;
                    active = true;
                    fresh = true;
                    progressed = true;
                }
                const uint64_t idle = (uint64_t)__popcll(idle_mask);
                wcur += idle < avail ? idle : avail;
            }
        }
        if (wdone && __ballot(active) == 0) break;

        const uint64_t td = t * (uint64_t)D;
        if (fresh) {
            // Attempt 0 of every slot at once; copy reads go out together.
#pragma unroll
            for (int kk = 0; kk < D; ++kk) {
                const Attempt at = attempt(a.key0, a.key1, T, t, (uint32_t)kk, 0u, td);
                if (at.direct)
                    val[kk] = (int32_t)(at.pair / (uint64_t)D);
                else
                    val[kk] = G[at.pair];
            }
            c_draws += D;
            k = 0;
            r = 0;
            need_draw = false;
            have_src = false;
            fresh = false;
        } else if (active) {
            // One redraw or one coherent re-read for the current slot k.
            int32_t got = kEmpty;
            bool resolved = false;
            if (need_draw) {
                const Attempt at = attempt(a.key0, a.key1, T, t, k, r, td);
                ++c_draws;
                need_draw = false;
                if (at.direct) {
                    got = (int32_t)(at.pair / (uint64_t)D);
                    resolved = true;
                    progressed = true;
                } else {
                    src = at.pair;
                    have_src = true;
                }
            }
            if (have_src) {
                got = load_coherent(G + src);
                ++c_polls;
                if (got != kEmpty) {
                    resolved = true;
                    progressed = true;
                    have_src = false;
                } else {
                    ++c_stalls;
                }
            }
            if (resolved) {
#pragma unroll
                for (int kk = 0; kk < D; ++kk)
                    if ((uint32_t)kk == k) val[kk] = got;
            }
        }

        // ---- accept slots in order: duplicate check against filled slots ----
        if (active && !need_draw && !have_src) {
            bool stop = false;
#pragma unroll
            for (int kk = 0; kk < D; ++kk) {
                if (!stop && (uint32_t)kk == k) {
                    const int32_t c = val[kk];
                    if (c == kEmpty) {
                        // attempt 0's copy source was still empty: re-derive it
                        need_draw = true;
                        stop = true;
                        ++c_pmiss;
                    } else {
                        bool dup = false;
#pragma unroll
                        for (int i = 0; i < kk; ++i) dup |= (val[i] == c);
                        if (dup) {
                            ++r;
                            ++c_retries;
                            need_draw = true;
                            stop = true;
                            if (r >= kMaxRetry) fail = kErrRetry;
                        } else {
                            ++k;
                            r = 0;
                            progressed = true;
                        }
                    }
                }
            }
            if (k == (uint32_t)D) {
                int32_t *row = G + td;
#pragma unroll
                for (int kk = 0; kk < D; ++kk) store_coherent(row + kk, val[kk]);
                active = false;
                ++c_nodes;
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
            ++c_sleeps;
            for (uint32_t s = 0; s < backoff; ++s) __builtin_amdgcn_s_sleep(4);
            backoff = backoff < 32 ? backoff * 2 : 32;
        } else {
            backoff = 1;
        }
    }

    // ---- statistics: one atomic per counter per wavefront ----
    const uint32_t sums[8] = {wave_sum(c_nodes), wave_sum(c_draws), wave_sum(c_retries),
                              wave_sum(c_pmiss), wave_sum(c_polls), wave_sum(c_stalls),
                              wave_sum(c_claims), wave_sum(c_sleeps)};
    if (lane == 0) {
        unsigned long long *st = a.scratch->stats;
        atomicAdd(&st[kStatNodes], (unsigned long long)sums[0]);
        atomicAdd(&st[kStatSlots], (unsigned long long)sums[0] * D);
        atomicAdd(&st[kStatDraws], (unsigned long long)sums[1]);
        atomicAdd(&st[kStatRetries], (unsigned long long)sums[2]);
        atomicAdd(&st[kStatPrefetchMisses], (unsigned long long)sums[3]);
        atomicAdd(&st[kStatPolls], (unsigned long long)sums[4]);
        atomicAdd(&st[kStatStalls], (unsigned long long)sums[5]);
        atomicAdd(&st[kStatClaims], (unsigned long long)sums[6]);
        atomicAdd(&st[kStatSleeps], (unsigned long long)sums[7]);
    }
}

// ---- degree pass ----------------------------------------------------------

__global__ void degree_init_kernel(uint32_t *deg, uint64_t n, uint32_t d)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) deg[v] = d;
}

// deg[target]++ for every slot of nodes t >= d+1 (cuPPA.cu:115-131).
// Out-of-range values (an incomplete graph) are skipped, never dereferenced.
__global__ void degree_count_kernel(const int32_t *G, uint64_t first_slot, uint64_t slots, uint64_t n,
                                    uint32_t *deg)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t s = first_slot + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < slots; s += stride) {
        const int32_t v = G[s];
        if (v >= 0 && (uint64_t)v < n) atomicAdd(&deg[v], 1u);
    }
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
            if (x < lds_len)
                atomicAdd(&local[x], 1u);
            else if (x < hist_len)
                atomicAdd((unsigned long long *)&hist[x], 1ull);
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

template <int D>
hipError_t launch_generate_d(const GenArgs &a, uint32_t blocks, uint32_t threads, hipStream_t stream)
{
    hipLaunchKernelGGL(generate_kernel<D>, dim3(blocks), dim3(threads), 0, stream, a);
    return hipGetLastError();
}

template <int D>
hipError_t occupancy_d(uint32_t threads, int *out)
{
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(out, generate_kernel<D>, (int)threads, 0);
}

}  // namespace

hipError_t launch_reset_scratch(DeviceScratch *s, uint64_t first_node, hipStream_t stream)
{
    hipLaunchKernelGGL(reset_scratch_kernel, dim3(1), dim3(64), 0, stream, s, first_node);
    return hipGetLastError();
}

hipError_t launch_fill(const FillArgs &a, hipStream_t stream)
{
    const uint64_t work = (a.n * a.d) / 4 + 1;
    hipLaunchKernelGGL(fill_kernel, dim3(grid_for(work, 256, 8192)), dim3(256), 0, stream, a);
    return hipGetLastError();
}

#define PPA_FOR_EACH_D(X) \
    X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16)

hipError_t launch_generate(uint32_t d, const GenArgs &a, uint32_t blocks, uint32_t threads,
                           hipStream_t stream)
{
    switch (d) {
#define PPA_CASE(DD) \
    case DD: return launch_generate_d<DD>(a, blocks, threads, stream);
        PPA_FOR_EACH_D(PPA_CASE)
#undef PPA_CASE
    default: return hipErrorInvalidValue;
    }
}

hipError_t generate_occupancy(uint32_t d, uint32_t threads, int *out)
{
    switch (d) {
#define PPA_CASE(DD) \
    case DD: return occupancy_d<DD>(threads, out);
        PPA_FOR_EACH_D(PPA_CASE)
#undef PPA_CASE
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_degree(const int32_t *G, uint64_t n, uint32_t d, uint32_t *deg, uint64_t *hist,
                         uint64_t hist_len, DeviceScratch *s, hipStream_t stream)
{
    hipLaunchKernelGGL(degree_init_kernel, dim3(grid_for(n, 256, 8192)), dim3(256), 0, stream, deg, n, d);
    const uint64_t first = (uint64_t)(d + 1) * d, slots = n * d;
    if (slots > first)
        hipLaunchKernelGGL(degree_count_kernel, dim3(grid_for(slots - first, 256, 16384)), dim3(256), 0,
                           stream, G, first, slots, n, deg);
    if (hist && hist_len)
        hipLaunchKernelGGL(zero_u64_kernel, dim3(grid_for(hist_len, 256, 4096)), dim3(256), 0, stream, hist,
                           hist_len);
    hipLaunchKernelGGL(histogram_kernel, dim3(grid_for(n, 256, 2048)), dim3(256), 0, stream, deg, n, hist,
                       hist_len, s);
    return hipGetLastError();
}

}  // namespace ppa
