// ppa_device.hpp -- wavefront helpers shared by the gfx950 generation kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ppa {

constexpr int kWave = 64;
constexpr uint32_t kMaxRetry = 1u << 24;
constexpr int kErrRetry = -2;
constexpr int kErrTimeout = -5;
constexpr int kErrInternal = -10;
constexpr int kErrRange = -11;

// A 32-bit count that would pass 2^32 - 1 when `add` is added to `old` (the
// value before an atomic add) is reported through *status, first error wins.
__device__ __forceinline__ void check_u32_add(uint32_t old, uint32_t add, int32_t *status)
{
    if (old > 0xFFFFFFFFu - add) atomicCAS(status, 0, kErrRange);
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// The value of lane 0 (an SGPR), for values every lane already agrees on.
__device__ __forceinline__ uint64_t uniform64(uint64_t v)
{
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// Number of set bits of `mask` below this lane.
__device__ __forceinline__ uint32_t rank_below(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Agent-scope relaxed accesses: `sc1` loads and stores on gfx950, coherent
// across the XCDs' separate L2 caches.
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

__device__ __forceinline__ uint64_t wave_sum64(uint64_t x)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t lo = __shfl_xor((uint32_t)x, off, kWave);
        const uint32_t hi = __shfl_xor((uint32_t)(x >> 32), off, kWave);
        x += ((uint64_t)hi << 32) | lo;
    }
    return x;
}

// floor(R / d) for a runtime divisor 1 <= d <= 2^16 and R < 2^52: a double
// estimate off by at most one, corrected with an exact integer remainder.
__device__ __forceinline__ uint64_t div_small(uint64_t R, uint32_t d, double inv_d)
{
    uint64_t q = (uint64_t)((double)R * inv_d);
    const int64_t rem = (int64_t)(R - q * d);
    if (rem < 0) --q;
    else if (rem >= (int64_t)d) ++q;
    return q;
}

}  // namespace ppa
