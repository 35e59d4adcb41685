// ppa_probe.hip -- measurement aid for bench.py, not part of include/ppa.h.
//
// The generator's copy reads are uniform random 4-byte loads over the target
// array, each one 64-byte HBM request. Their ceiling on this GPU is therefore
// a random-request rate, not a byte rate: this kernel measures it with the
// same kind of load. Every lane issues `iters` loads at hashed positions
// below table_len, eight in flight at a time, and writes one sum to `sink`
// (blocks * threads entries) so no load is dead. Positions are
// mulhi64(hash, table_len) < table_len: every access is in bounds.
#include <hip/hip_runtime.h>
#include <stdint.h>

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

}  // namespace

extern "C" {

// Times one launch of blocks x threads lanes, `iters` (a multiple of 8)
// random loads each, on `stream`; *ms_out receives the launch's duration.
// Returns 0, or -1 for bad arguments, or -4 for a HIP error.
int ppa_probe_gather(const int32_t *table, uint64_t table_len, uint32_t blocks, uint32_t threads, uint32_t iters,
                     int nontemporal, int32_t *sink, void *stream, float *ms_out)
{
    if (!table || !sink || !ms_out || table_len == 0 || blocks == 0 || threads == 0 || threads > 256 ||
        threads % 64 != 0 || iters == 0 || iters % 8 != 0)
        return -1;
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return -4;
    hipError_t err = hipEventRecord(e0, s);
    if (nontemporal)
        hipLaunchKernelGGL(gather_kernel<true>, dim3(blocks), dim3(threads), 0, s, table, table_len, iters,
                           0x9e3779b97f4a7c15ull, sink);
    else
        hipLaunchKernelGGL(gather_kernel<false>, dim3(blocks), dim3(threads), 0, s, table, table_len, iters,
                           0x9e3779b97f4a7c15ull, sink);
    if (err == hipSuccess) err = hipGetLastError();
    if (err == hipSuccess) err = hipEventRecord(e1, s);
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    if (err == hipSuccess) err = hipEventElapsedTime(ms_out, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return err == hipSuccess ? 0 : -4;
}

}  // extern "C"
