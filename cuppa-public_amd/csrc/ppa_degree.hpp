// ppa_degree.hpp -- the partitioned degree pass (ppa_degree.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ppa {

// Device buffers of one partitioned degree pass, carved from one allocation
// by degree_work_layout (the in-edge CSR borrows counts, offsets and
// partials for its radix passes).
struct DegreeWork {
    int32_t *part_a = nullptr;      // slots after the level-1 partition (graphs of more than 2^23 nodes)
    uint16_t *part_b = nullptr;     // slots in leaf order, low 15 bits of each target
    uint32_t *c1 = nullptr;         // level-1 counts per (bucket, workgroup)
    uint64_t *offs1 = nullptr;      // their exclusive scan, one extra entry for the total
    uint32_t *c2 = nullptr;         // level-2 counts per (bucket, leaf, workgroup)
    uint64_t *offs2 = nullptr;      // their exclusive scan, one extra entry for the total
    uint64_t *partials = nullptr;   // scan workgroup sums
    uint64_t *whole = nullptr;      // {0, slots}: the one leaf of a graph of at most 2^15 nodes
    uint64_t *piece_base = nullptr; // exclusive scan of pieces per leaf
    uint32_t *hub = nullptr;        // in-slot counts of the hub nodes, from the count pass
    int32_t *status = nullptr;      // set to -11 (PPA_ERR_RANGE) if a degree passes 2^32 - 1
    uint64_t *counts = nullptr;     // in-edge CSR: per (digit, workgroup) counts
    uint64_t *offsets = nullptr;    // in-edge CSR: their exclusive scan
    uint32_t num_cus = 0;           // compute units the launch shapes are sized for
};

// Bytes of the DegreeWork of an n-node, out-degree-d graph on a GPU with
// num_cus compute units.
size_t degree_work_bytes(uint64_t n, uint32_t d, uint32_t num_cus);
void degree_work_layout(void *base, uint64_t n, uint32_t d, uint32_t num_cus, DegreeWork *w);

// deg[v] = d + #{slots of nodes t >= d+1 naming v} for v < n.
hipError_t launch_degree_partitioned(const int32_t *G, uint64_t n, uint32_t d, uint32_t *deg, const DegreeWork &w,
                                     hipStream_t stream);

// In-edge CSR of all n*d slots (seed rows included): in_src[in_ptr[v] ..
// in_ptr[v+1]) are the sources of the slots naming v in ascending order;
// in_ptr[n] = valid slots. work holds csr_work_bytes(n, d, num_cus) bytes.
size_t csr_work_bytes(uint64_t n, uint32_t d, uint32_t num_cus);
// sorted_keys (nullable) receives the work-area buffer holding the targets in
// in_src order, valid until the work area is reused.
hipError_t launch_in_csr(const int32_t *G, uint64_t n, uint32_t d, int64_t *in_ptr, int32_t *in_src, void *work,
                         uint32_t num_cus, hipStream_t stream, const int32_t **sorted_keys = nullptr);
// Undirected lists (out row, then the in-list without seed-row sources) of a
// complete graph, given its in-edge CSR and the sorted targets of launch_in_csr.
// ptr int64[n+1], adj int32[ptr[n]].
hipError_t launch_undirected(const int32_t *G, uint64_t n, uint32_t d, const int64_t *in_ptr, const int32_t *keys,
                             const int32_t *in_src, int64_t *ptr, int32_t *adj, hipStream_t stream);

}  // namespace ppa
