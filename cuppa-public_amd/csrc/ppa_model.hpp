// ppa_model.hpp -- keyed draws of the copy model, shared by host and device.
//
// The reference draws with a per-thread cuRAND XORWOW stream seeded from the
// wall clock (cuRandom.cu:47-52, cuPPA_1.cu:39-40): `curand % u` for the node,
// `curand_uniform` against p for the coin, `curand % M` for the copied slot
// (cuPPA_RRP_1.cu:70-85). Its serial path draws one 64-bit value over the
// M*t (node, slot) pairs and a double coin (SerialCopyModel3.cpp:62-67).
// Here every attempt is one stateless Philox-4x32-10 call keyed by
// (seed; t, k, r), written from Salmon et al., "Parallel random numbers: as
// easy as 1, 2, 3" (SC'11): x0 is the coin, (x1, x2) the 64-bit (node, slot)
// draw, which keeps the serial path's single draw over the t*d pairs.
#pragma once

#include <stdint.h>

#include "ppa.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PPA_HD __host__ __device__ __forceinline__
#else
#define PPA_HD inline
#endif

namespace ppa {

constexpr int32_t kEmpty = -1;  // an unfilled slot (SerialCopyModel.cpp:41 uses -1 too)

struct Draw {
    uint32_t x0, x1, x2, x3;
};

PPA_HD uint32_t mul_hi32(uint32_t a, uint32_t b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __umulhi(a, b);
#else
    return (uint32_t)(((uint64_t)a * b) >> 32);
#endif
}

// Philox-4x32-10: ten rounds of the 4x32 S-box with key schedule
// (k0, k1) += (0x9E3779B9, 0xBB67AE85) between rounds.
PPA_HD Draw philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        // one 32x32->64 product per multiplier (v_mad_u64_u32 on gfx950)
        const uint64_t pa = (uint64_t)0xD2511F53u * c0;
        const uint64_t pb = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t lo_a = (uint32_t)pa, hi_a = (uint32_t)(pa >> 32);
        const uint32_t lo_b = (uint32_t)pb, hi_b = (uint32_t)(pb >> 32);
        const uint32_t n0 = hi_b ^ c1 ^ k0;
        const uint32_t n2 = hi_a ^ c3 ^ k1;
        c0 = n0;
        c1 = lo_b;
        c2 = n2;
        c3 = lo_a;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return Draw{c0, c1, c2, c3};
}

PPA_HD uint64_t mul_hi64(uint64_t a, uint64_t b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(a, b);
#else
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

// Attempt r of slot k of node t: the coin and the (node, slot) pair.
struct Attempt {
    bool direct;
    uint64_t pair;  // R = u*d + j, uniform over [0, t*d)
};

PPA_HD Attempt attempt(uint32_t key0, uint32_t key1, uint64_t T, uint64_t t, uint32_t k, uint32_t r,
                       uint64_t td)
{
    const Draw x = philox((uint32_t)t, (uint32_t)(t >> 32), k, r, key0, key1);
    Attempt a;
    a.direct = (uint64_t)x.x0 < T;
    a.pair = mul_hi64(((uint64_t)x.x1 << 32) | x.x2, td);
    return a;
}

// Seed clique on nodes 0..d: node i's slot j names j + (j >= i).
PPA_HD int32_t clique_target(uint64_t i, uint32_t j) { return (int32_t)(j + (j >= i ? 1u : 0u)); }

// Seed rows 0..d of a graph of the given model. Barabasi-Albert seed nodes
// i < d hold self-loops (SerialBA.cpp:90-97); its node d names 0..d-1
// (SerialBA.cpp:108-112), which is the clique's row d.
PPA_HD int32_t seed_target(uint32_t model, uint32_t d, uint64_t i, uint32_t j)
{
    return model == PPA_MODEL_BA && i < d ? (int32_t)i : clique_target(i, j);
}

// Coin threshold the generator uses: a Barabasi-Albert endpoint is the source
// of its (node, slot) pair with probability exactly 1/2.
PPA_HD uint64_t coin_threshold(uint32_t model, uint64_t p_threshold)
{
    return model == PPA_MODEL_BA ? (1ull << 31) : p_threshold;
}

// Degree every node starts from before the slots of nodes t >= d+1 are
// counted: its own d slots (copy model), or, for Barabasi-Albert, d for
// v >= d and the one slot of node d for a seed node (self-loops uncounted).
PPA_HD uint32_t base_degree(uint32_t model, uint32_t d, uint64_t v)
{
    return model == PPA_MODEL_BA && v < d ? 1u : d;
}

}  // namespace ppa
