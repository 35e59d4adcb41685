// ppa_probe.hip -- measurement aid for bench.py, not part of include/ppa.h.
//
// The generator's copy reads are uniform random 4-byte loads over the target
// array, each one 64-byte HBM request. Their ceiling on this GPU is therefore
// a random-request rate, not a byte rate: these kernels measure it with the
// same kind of load, into registers or, as the generator's attempt-0 reads
// do, straight into LDS (global_load_lds). Every lane issues `iters` loads at
// hashed positions below table_len, eight in flight at a time, and writes one
// sum to `sink` (blocks * threads entries) so no load is dead. Positions are
// mulhi64(hash, table_len) < table_len: every access is in bounds.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ppa_model.hpp"

namespace {

__device__ __forceinline__ uint64_t mix64(uint64_t x)
{
    // SplitMix64 finaliser (Steele, Lea and Flood, OOPSLA 2014)
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

template <bool NT>
__global__ __launch_bounds__(256) void gather_kernel(const int32_t *table, uint64_t table_len, uint32_t iters,
                                                     uint64_t seed, int32_t *sink)
{
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int32_t acc = 0;
#pragma unroll 1
    for (uint32_t i = 0; i < iters; i += 8) {
        int32_t v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const uint64_t pos = __umul64hi(mix64(seed ^ (tid << 20) ^ (uint64_t)(i + u)), table_len);
            v[u] = NT ? __builtin_nontemporal_load(table + pos) : table[pos];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += v[u];
    }
    sink[tid] = acc;
}

// The same loads into LDS: eight 64-lane rows per wavefront, one wait per
// batch of eight, then each lane sums its own words.
template <bool NT>
__global__ __launch_bounds__(256) void gather_lds_kernel(const int32_t *table, uint64_t table_len, uint32_t iters,
                                                         uint64_t seed, int32_t *sink)
{
    __shared__ int32_t rows[4][8 * 64];
    int32_t *const mine = rows[threadIdx.x >> 6];
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int32_t acc = 0;
#pragma unroll 1
    for (uint32_t i = 0; i < iters; i += 8) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const uint64_t pos = __umul64hi(mix64(seed ^ (tid << 20) ^ (uint64_t)(i + u)), table_len);
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)(table + pos),
                                             (__attribute__((address_space(3))) void *)(mine + u * 64), 4, 0,
                                             NT ? 2 : 0);
        }
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += mine[u * 64 + lane];
        __builtin_amdgcn_wave_barrier();
    }
    sink[tid] = acc;
}

// gather_lds_kernel with the load's cache-policy bits spelled out (gfx950:
// sc0 = 1, nt = 2, sc1 = 16): which form of an isolated 4-byte read reaches
// memory as the smallest request, and at what rate.
template <int AUX>
__global__ __launch_bounds__(256) void gather_policy_kernel(const int32_t *table, uint64_t table_len, uint32_t iters,
                                                            uint64_t seed, int32_t *sink)
{
    __shared__ int32_t rows[4][8 * 64];
    int32_t *const mine = rows[threadIdx.x >> 6];
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int32_t acc = 0;
#pragma unroll 1
    for (uint32_t i = 0; i < iters; i += 8) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const uint64_t pos = __umul64hi(mix64(seed ^ (tid << 20) ^ (uint64_t)(i + u)), table_len);
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)(table + pos),
                                             (__attribute__((address_space(3))) void *)(mine + u * 64), 4, 0, AUX);
        }
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += mine[u * 64 + lane];
        __builtin_amdgcn_wave_barrier();
    }
    sink[tid] = acc;
}

// The generator's draw alone: `iters` keyed Philox-4x32-10 attempts per lane
// (the coin and the mulhi64 pair mapping included), no memory traffic but
// one store per lane. Its rate is the integer-VALU roof of the draws.
__global__ __launch_bounds__(256) void philox_kernel(uint32_t iters, uint64_t seed, int32_t *sink)
{
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    uint64_t acc = 0;
    const uint64_t t = (1ull << 29) + tid;
#pragma unroll 1
    for (uint32_t i = 0; i < iters; ++i) {
        const ppa::Attempt a = ppa::attempt(k0, k1, 858993459ull, t, i & 7u, i >> 3, t * 8u);
        acc += a.pair + (a.direct ? 1u : 0u);
    }
    sink[tid] = (int32_t)(acc ^ (acc >> 32));
}

typedef unsigned v4u __attribute__((ext_vector_type(4)));

// gather_lds_kernel<true> plus, per batch, the row publication the generator
// does per pass: each wavefront stores one contiguous 2 KB block (two 16-byte
// agent-scope stores per lane) into the first `rows_len` ints of the table,
// block after block, before waiting for everything.
__global__ __launch_bounds__(256) void gather_store_kernel(int32_t *table, uint64_t table_len, uint64_t rows_len,
                                                           uint32_t iters, uint64_t seed, int32_t *sink)
{
    __shared__ int32_t rows[4][8 * 64];
    int32_t *const mine = rows[threadIdx.x >> 6];
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t wave = tid >> 6, nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const uint64_t blocks = rows_len / 512;  // 2 KB blocks
    int32_t acc = 0;
#pragma unroll 1
    for (uint32_t i = 0; i < iters; i += 8) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const uint64_t pos = __umul64hi(mix64(seed ^ (tid << 20) ^ (uint64_t)(i + u)), table_len);
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)(table + pos),
                                             (__attribute__((address_space(3))) void *)(mine + u * 64), 4, 0, 2);
        }
        const uint64_t blk = (wave + (uint64_t)(i / 8) * nwaves) % blocks;
        const __amdgpu_buffer_rsrc_t out = __builtin_amdgcn_make_buffer_rsrc(table + blk * 512, 0, 2048, 0x00020000);
        const v4u v = {(unsigned)acc, (unsigned)i, (unsigned)lane, 1u};
        __builtin_amdgcn_raw_buffer_store_b128(v, out, (int)(lane * 16u), 0, 16);
        __builtin_amdgcn_raw_buffer_store_b128(v, out, (int)(1024u + lane * 16u), 0, 16);
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += mine[u * 64 + lane];
        __builtin_amdgcn_wave_barrier();
    }
    sink[tid] = acc;
}

}  // namespace

extern "C" {

// Times one launch of blocks x threads lanes, `iters` (a multiple of 8)
// random loads each, on `stream`; *ms_out receives the launch's duration.
// mode: 0 plain loads, 1 non-temporal, 2 into LDS, 3 into LDS non-temporal,
// 4 as 3 with one 2 KB block of agent-scope row stores per wavefront and
// batch (written into the table's first 2^28 ints: table_len >= 2^28): the
// generator's mix of copy reads and row publication; 5..10 into LDS with the
// cache-policy bits sc0, sc1, sc0 sc1, nt sc0, nt sc1, nt sc0 sc1; 11 the
// Philox draws alone (`iters` per lane; table and table_len unused).
// Returns 0, or -1 for bad arguments, or -4 for a HIP error.
int ppa_probe_gather(const int32_t *table, uint64_t table_len, uint32_t blocks, uint32_t threads, uint32_t iters,
                     int mode, int32_t *sink, void *stream, float *ms_out)
{
    if (!table || !sink || !ms_out || table_len == 0 || blocks == 0 || threads == 0 || threads > 256 ||
        threads % 64 != 0 || iters == 0 || iters % 8 != 0 || mode < 0 || mode > 11 ||
        (mode == 4 && table_len < (1ull << 28)))
        return -1;
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return -4;
    hipError_t err = hipEventRecord(e0, s);
    const uint64_t seed = 0x9e3779b97f4a7c15ull;
    if (mode == 0)
        hipLaunchKernelGGL(gather_kernel<false>, dim3(blocks), dim3(threads), 0, s, table, table_len, iters, seed, sink);
    else if (mode == 1)
        hipLaunchKernelGGL(gather_kernel<true>, dim3(blocks), dim3(threads), 0, s, table, table_len, iters, seed, sink);
    else if (mode == 2)
        hipLaunchKernelGGL(gather_lds_kernel<false>, dim3(blocks), dim3(threads), 0, s, table, table_len, iters, seed,
                           sink);
    else if (mode == 3)
        hipLaunchKernelGGL(gather_lds_kernel<true>, dim3(blocks), dim3(threads), 0, s, table, table_len, iters, seed,
                           sink);
    else if (mode == 4)
        hipLaunchKernelGGL(gather_store_kernel, dim3(blocks), dim3(threads), 0, s, const_cast<int32_t *>(table),
                           table_len, 1ull << 28, iters, seed, sink);
    else if (mode == 5)
        hipLaunchKernelGGL(gather_policy_kernel<1>, dim3(blocks), dim3(threads), 0, s, table, table_len, iters, seed,
                           sink);
    else if (mode == 6)
        hipLaunchKernelGGL(gather_policy_kernel<16>, dim3(blocks), dim3(threads), 0, s, table, table_len, iters, seed,
                           sink);
    else if (mode == 7)
        hipLaunchKernelGGL(gather_policy_kernel<17>, dim3(blocks), dim3(threads), 0, s, table, table_len, iters, seed,
                           sink);
    else if (mode == 8)
        hipLaunchKernelGGL(gather_policy_kernel<3>, dim3(blocks), dim3(threads), 0, s, table, table_len, iters, seed,
                           sink);
    else if (mode == 9)
        hipLaunchKernelGGL(gather_policy_kernel<18>, dim3(blocks), dim3(threads), 0, s, table, table_len, iters, seed,
                           sink);
    else if (mode == 10)
        hipLaunchKernelGGL(gather_policy_kernel<19>, dim3(blocks), dim3(threads), 0, s, table, table_len, iters, seed,
                           sink);
    else
        hipLaunchKernelGGL(philox_kernel, dim3(blocks), dim3(threads), 0, s, iters, seed, sink);
    if (err == hipSuccess) err = hipGetLastError();
    if (err == hipSuccess) err = hipEventRecord(e1, s);
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    if (err == hipSuccess) err = hipEventElapsedTime(ms_out, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return err == hipSuccess ? 0 : -4;
}

}  // extern "C"
