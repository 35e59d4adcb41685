// ppa_noalloc.cpp -- ppa_generate for callers that cannot allocate device
// memory (tight serving loops, harnesses that forbid hipMalloc, code that must
// not allocate around stream captures; it waits for its stream, so the call
// itself is not capturable). It has no counterpart in the reference.
//
// The generator's only device memory beyond the caller's graph is one
// DeviceScratch record per launch: the claim counter, the status word and the
// counters. Here that record lives in `scratch`, a file-scope static device
// array of kStaticScratchBytes that every GPU the library runs on holds until
// the process ends. In host code the array's name is only a stand-in for the
// symbol: its device address is looked up once per device, with that device
// current, through hipGetSymbolAddress, and only that address is handed to
// kernels. Each launch resets the record with a kernel before the generator
// runs (no memset node), so a call never sees the previous call's contents.
// Calls that overlap on one GPU share the record and race.
#include <cstring>
#include <mutex>

#include "ppa_internal.hpp"

namespace {

using namespace ppa;

constexpr size_t kStaticScratchBytes = 256;
static_assert(sizeof(DeviceScratch) <= kStaticScratchBytes, "one launch record fits the static scratch");

static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

constexpr int kMaxDevices = 64;

struct DeviceSlot {
    std::once_flag once;
    int rc = PPA_OK;
    DeviceScratch *ptr = nullptr;  // this device's copy of `scratch`
    int num_cus = 0;
    int wall_clock_khz = 0;
};

DeviceSlot g_slots[kMaxDevices];

// Once per device, with it current: the device address of `scratch` and the
// attributes the launch shape needs.
void set_up(DeviceSlot &s, int dev)
{
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) {
        (void)hipGetLastError();
        s.rc = PPA_ERR_DEVICE;
        return;
    }
    if (hipDeviceGetAttribute(&s.num_cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipDeviceGetAttribute(&s.wall_clock_khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess) {
        s.rc = PPA_ERR_DEVICE;
        return;
    }
    if (s.wall_clock_khz <= 0) s.wall_clock_khz = 100000;
    s.ptr = static_cast<DeviceScratch *>(p);
    s.rc = PPA_OK;
}

}  // namespace

extern "C" int ppa_generate_noalloc(const ppa_graph_desc *desc, int32_t *d_targets, int64_t *d_row_ptr,
                                    ppa_stats *stats, void *stream_)
{
    if (!d_targets) return PPA_ERR_INVALID;
    int rc = ppa_check_desc(desc);
    if (rc) return rc;
    if (((uintptr_t)d_targets & 15u) != 0) return PPA_ERR_INVALID;
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) return PPA_ERR_DEVICE;
    if (dev >= kMaxDevices) return PPA_ERR_INVALID;
    DeviceSlot &slot = g_slots[dev];
    std::call_once(slot.once, set_up, std::ref(slot), dev);
    if (slot.rc != PPA_OK) return slot.rc;

    // The launch rules of ppa_generate with the default tuning; a stack
    // context carries the device's attributes (its members allocate nothing).
    ppa_ctx c;
    c.device = dev;
    c.num_cus = slot.num_cus;
    c.wall_clock_khz = slot.wall_clock_khz;
    const uint64_t n = desc->n;
    const uint32_t d = desc->d;
    const uint64_t start = single_start(&c, *desc);
    LaunchShape shape;
    if ((rc = resolve_shape(&c, d, n > start ? n - start : 1, &shape)) != PPA_OK) return rc;
    hipStream_t stream = (hipStream_t)stream_;
    PPA_HIP(enqueue_single(&c, *desc, d_targets, d_row_ptr, shape, slot.ptr, stream));
    PPA_HIP(hipStreamSynchronize(stream));
    DeviceScratch host;
    std::memset(&host, 0, sizeof(host));
    PPA_HIP(hipMemcpy(&host, slot.ptr, sizeof(host), hipMemcpyDeviceToHost));
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->nodes = host.stats[kStatNodes];
        stats->slots = host.stats[kStatSlots];
        stats->draws = host.stats[kStatDraws];
        stats->retries = host.stats[kStatRetries];
        stats->prefetch_misses = host.stats[kStatPrefetchMisses];
        stats->polls = host.stats[kStatPolls];
        stats->stalls = host.stats[kStatStalls];
        stats->claims = host.stats[kStatClaims];
        stats->sleeps = host.stats[kStatSleeps];
        stats->staged_rows = host.stats[kStatStagedRows];
        stats->passes = host.stats[kStatPasses];
        stats->grid_blocks = shape.blocks;
        stats->block_threads = shape.threads;
    }
    if (host.status != 0) return host.status;
    if (host.stats[kStatNodes] != n - d - 1) return PPA_ERR_INTERNAL;
    return PPA_OK;
}
