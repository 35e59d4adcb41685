// ppa_internal.hpp -- context state shared by the C ABI implementation files.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/ppa.h"
#include "ppa_kernels.hpp"
#include "ppa_model.hpp"

struct ppa_ctx {
    int device = 0;
    int num_cus = 0;
    int wall_clock_khz = 0;
    ppa_tuning tuning{};
    ppa::DeviceScratch *scratch = nullptr;  // device
    uint32_t *degree_scratch = nullptr;     // device, grown on demand
    uint64_t degree_scratch_len = 0;
    void *work = nullptr;                   // device work area of the partition passes (degree, in-edge CSR)
    size_t work_bytes = 0;
    uint64_t scratch_budget = 0;            // cap on work + degree_scratch bytes (0 = none)
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    ppa_stats last{};
    bool pending = false;        // a launch awaits ppa_ctx_sync / ppa_batch_sync
    uint64_t expect_nodes = 0;   // generated nodes the pending launch must finish
    // ppa_generate_range: the graph being generated range by range, and the
    // node its next range must start at (0: none in progress)
    ppa_graph_desc range_desc{};
    int32_t *range_targets = nullptr;
    uint64_t range_next = 0;
    std::map<std::string, ppa_tuning> tune_cache;
};

namespace ppa {

constexpr uint32_t kDefaultBlock = 256;
constexpr uint32_t kDefaultChunk = 256;
constexpr uint32_t kDefaultWatchdogMs = 60000;
constexpr uint32_t kDefaultEarlyNodes = 1u << 22;
constexpr uint32_t kEarlyChunk = 64;
// Single-graph generator launches hold at most this many wavefronts per CU
// unless the tuning's blocks_per_cu says otherwise.
constexpr uint32_t kSingleWavesPerCu = 16;  // n = 2^30, d = 8: 16 waves 174.2 ms, 20: 175.0, 24: 176.0, 32: 178.3

class DeviceGuard {
public:
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&prev_) != hipSuccess) prev_ = -1;
        ok_ = hipSetDevice(dev) == hipSuccess;
    }
    ~DeviceGuard()
    {
        if (prev_ >= 0) (void)hipSetDevice(prev_);
    }
    bool ok() const { return ok_; }

private:
    int prev_ = -1;
    bool ok_ = false;
};

#define PPA_HIP(call)                                    \
    do {                                                 \
        if ((call) != hipSuccess) return PPA_ERR_DEVICE; \
    } while (0)

struct LaunchShape {
    uint32_t blocks = 0, threads = 0, chunk = 0;
    uint64_t early_nodes = 0;
    uint64_t watchdog_ticks = 0;
    uint32_t sleep_cap = 1;
    uint32_t wide_mask = 3;
};

// Grid, workgroup and claim sizes for `work` generated nodes with out-degree
// d (d = 0: the batch kernel).
// cap_single: hold single graphs (d != 0) to kSingleWavesPerCu per CU.
int resolve_shape(const ppa_ctx *ctx, uint32_t d, uint64_t work, LaunchShape *out, bool cap_single = true);
// First node the persistent generator claims for a single graph: after the
// LDS-resident prefix, or d + 1 without it.
uint64_t single_start(const ppa_ctx *ctx, const ppa_graph_desc &desc);
// Reset, fill, prefix and persistent generator of one graph on `stream`;
// `shape` was resolved for the work range_start(...) .. last. after_fill
// (nullable) is recorded between the fill and the generation. A node range
// [first, last) (last = 0: n) generates only those nodes: first = 0 fills the
// whole graph and runs the prefix (cut at `last`); first > 0 needs nodes below
// it final and the graph filled.
hipError_t enqueue_single(const ppa_ctx *ctx, const ppa_graph_desc &desc, int32_t *G, int64_t *row_ptr,
                          const LaunchShape &shape, DeviceScratch *scratch, hipStream_t stream,
                          hipEvent_t after_fill = nullptr, uint64_t first = 0, uint64_t last = 0);
// The first node the persistent generator claims in the range [first, last).
uint64_t range_start(const ppa_ctx *ctx, const ppa_graph_desc &desc, uint64_t first, uint64_t last);

// Grows the context's work area to `need` bytes (waiting for `stream` before
// freeing the old one); PPA_ERR_BUDGET if that would hold more than the
// context's scratch budget, PPA_ERR_NOMEM if the device cannot provide it.
int ensure_work(ppa_ctx *ctx, size_t need, hipStream_t stream);
// Bytes of work memory the context holds (work area + degree scratch).
uint64_t held_bytes(const ppa_ctx *ctx);

// Host image of a batch's device tables, packed as
// BatchGraph[count] | slot_base[count+1] | row_base[count+1] | seg_base[nseg+1] |
// seg_round[nseg] | seg_active[nseg] | order[count] (uint32).
struct BatchTables {
    std::vector<char> bytes;
    uint64_t count = 0;
    uint64_t nseg = 0;
    uint64_t work = 0;         // nodes for the persistent generator
    uint64_t nodes = 0;        // generated nodes, the prefixes' included
    uint64_t nprefix = 0;      // graphs with an LDS-resident prefix
    uint64_t total_slots = 0;  // sum of n*d
    uint64_t total_rows = 0;   // sum of n+1 over graphs with a row_ptr
    bool any_row_ptr = false;
};
// Upper bound of the packed size for `count` graphs (nseg <= count).
size_t batch_tables_bytes(uint64_t count);
// graphs[i].G / row_ptr / targets_offset must be set; the rest is derived here.
void build_batch_tables(const std::vector<BatchGraph> &graphs, uint64_t chunk, BatchTables *out);
// reset + fill + generate of a batch whose tables are already at d_tables.
// done_flag (nullable, mapped host memory, one word per graph) is set to 1
// for each graph as soon as it is complete.
hipError_t enqueue_batch(void *d_tables, const BatchTables &t, const LaunchShape &shape, DeviceScratch *scratch,
                         hipStream_t stream, uint32_t *done_flag = nullptr);
// A descriptor -> the kernel's graph record (G, row_ptr and offsets set by the caller).
// prefix_cap: the tuning's prefix_nodes (see prefix_start).
BatchGraph graph_record(const ppa_graph_desc &desc, uint32_t prefix_cap);
// Waits for the pending launch (its end event ctx->ev[2]) and turns the
// device scratch into ctx->last; clears ctx->pending on every path.
int finish_launch(ppa_ctx *ctx);
bool tuning_valid(const ppa_tuning &t);

}  // namespace ppa
