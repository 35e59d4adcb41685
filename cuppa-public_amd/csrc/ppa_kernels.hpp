// ppa_kernels.hpp -- launch interface between the C ABI and the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ppa {

// Device scratch of one context (one per call in flight on that context).
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
    kStatCount
};

struct DeviceScratch {
    unsigned long long claim;  // next unclaimed node
    int status;                // 0 = running/ok, else a negative PPA_ERR_* code
    int pad;
    unsigned long long stats[kStatCount];
    unsigned long long max_degree;
};

struct GenArgs {
    int32_t *G;           // targets, row-major [n][d]
    uint64_t n;
    uint64_t T;           // coin threshold
    uint32_t key0, key1;  // Philox key
    uint32_t chunk;       // nodes claimed per wavefront claim
    uint32_t pad;
    uint64_t watchdog_ticks;  // wall-clock ticks a wavefront may run
    DeviceScratch *scratch;
};

struct FillArgs {
    int32_t *G;
    int64_t *row_ptr;  // nullable
    uint64_t n;
    uint32_t d;
    uint32_t pad;
};

// Kernels (defined in ppa_kernels.hip).
hipError_t launch_reset_scratch(DeviceScratch *s, uint64_t first_node, hipStream_t stream);
hipError_t launch_fill(const FillArgs &a, hipStream_t stream);
hipError_t launch_generate(uint32_t d, const GenArgs &a, uint32_t blocks, uint32_t block_threads,
                           hipStream_t stream);
hipError_t generate_occupancy(uint32_t d, uint32_t block_threads, int *blocks_per_cu);
hipError_t launch_degree(const int32_t *G, uint64_t n, uint32_t d, uint32_t *deg, uint64_t *hist,
                         uint64_t hist_len, DeviceScratch *s, hipStream_t stream);

}  // namespace ppa
