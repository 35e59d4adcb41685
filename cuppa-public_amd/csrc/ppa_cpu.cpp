// ppa_cpu.cpp -- the serial CPU copy model of the product (`-A scm`); also
// runs the Barabasi-Albert model (PPA_MODEL_BA) through the same loop.
//
// Counterpart of SerialCopyModel::generateGraphCopyModelBST
// (src/PA/SPA/SerialCopyModel3.cpp:11-107): one host thread walks the nodes in
// order and fills each node's slots with redraw-until-unique. It consumes the
// same keyed Philox draws as the GPU generator (ppa_model.hpp), so the two
// produce identical graphs. The duplicate set is the row itself for d <= 32
// and an open-addressing hash set above, instead of the reference's
// BSTVertex (BSTVertex.h:189-235).
#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/ppa.h"
#include "ppa_model.hpp"

namespace {

// A node's accepted targets for d > 32: linear probing over the smallest
// power of two >= 2d words (at most half full), cleared per node.
class RowSet {
public:
    explicit RowSet(uint32_t d)
    {
        while ((1u << bits_) < 2u * d) ++bits_;
        w_.assign((size_t)1 << bits_, ppa::kEmpty);
    }
    void clear() { std::fill(w_.begin(), w_.end(), ppa::kEmpty); }
    // false if v is already in the set
    bool insert(int32_t v)
    {
        const uint32_t mask = (1u << bits_) - 1u;
        for (uint32_t i = ((uint32_t)v * 0x9E3779B1u) >> (32u - bits_);; i = (i + 1u) & mask) {
            if (w_[i] == v) return false;
            if (w_[i] == ppa::kEmpty) {
                w_[i] = v;
                return true;
            }
        }
    }

private:
    uint32_t bits_ = 1;
    std::vector<int32_t> w_;
};

}  // namespace

extern "C" {

// Host-only utilities of the ABI (no HIP), kept here with the serial path so
// that this file builds with a plain host compiler (the sanitizer build).
const char *ppa_strerror(int code)
{
    switch (code) {
    case PPA_OK: return "ok";
    case PPA_ERR_INVALID: return "invalid argument";
    case PPA_ERR_RETRY: return "retry limit exceeded for a slot";
    case PPA_ERR_NOMEM: return "out of memory";
    case PPA_ERR_DEVICE: return "HIP runtime error";
    case PPA_ERR_TIMEOUT: return "kernel watchdog expired";
    case PPA_ERR_BUDGET: return "request exceeds the memory budget";
    case PPA_ERR_NOT_FOUND: return "unknown request";
    case PPA_ERR_CANCELLED: return "request cancelled";
    case PPA_ERR_AGAIN: return "not ready, try again";
    case PPA_ERR_INTERNAL: return "internal error";
    case PPA_ERR_RANGE: return "a result exceeds its output type (degree above 2^32 - 1)";
    default: return "unknown error";
    }
}

uint64_t ppa_p_threshold(double p)
{
    if (!(p > 0.0)) return 0;
    if (p >= 1.0) return 1ull << 32;
    return (uint64_t)(p * 4294967296.0);  // exact scaling, then floor
}

int ppa_check_desc(const ppa_graph_desc *desc)
{
    if (!desc) return PPA_ERR_INVALID;
    if (desc->d < 1 || desc->d > PPA_MAX_D) return PPA_ERR_INVALID;
    if (desc->n < (uint64_t)desc->d + 1 || desc->n > PPA_MAX_N) return PPA_ERR_INVALID;
    if (desc->p_threshold > (1ull << 32)) return PPA_ERR_INVALID;
    if (desc->model != PPA_MODEL_COPY && desc->model != PPA_MODEL_BA) return PPA_ERR_INVALID;
    return PPA_OK;
}

int ppa_cpu_generate(const ppa_graph_desc *desc, int32_t *targets, int64_t *row_ptr, uint64_t *retries_out)
{
    int rc = ppa_check_desc(desc);
    if (rc) return rc;
    if (!targets) return PPA_ERR_INVALID;
    const uint64_t n = desc->n;
    const uint32_t d = desc->d;
    const uint64_t T = ppa::coin_threshold(desc->model, desc->p_threshold);
    const uint32_t k0 = (uint32_t)desc->seed, k1 = (uint32_t)(desc->seed >> 32);
    for (uint64_t i = 0; i <= d; ++i)
        for (uint32_t j = 0; j < d; ++j) targets[i * d + j] = ppa::seed_target(desc->model, d, i, j);
    RowSet *set = nullptr;
    if (d > 32) {
        set = new (std::nothrow) RowSet(d);
        if (!set) return PPA_ERR_NOMEM;
    }
    uint64_t retries = 0;
    for (uint64_t t = (uint64_t)d + 1; t < n; ++t) {
        int32_t *row = targets + t * d;
        const uint64_t td = t * d;
        if (set) set->clear();
        for (uint32_t k = 0; k < d; ++k) {
            for (uint32_t r = 0;; ++r) {
                if (r >= PPA_MAX_RETRY) {
                    delete set;
                    return PPA_ERR_RETRY;
                }
                const ppa::Attempt a = ppa::attempt(k0, k1, T, t, k, r, td);
                const int32_t cand = a.direct ? (int32_t)(a.pair / d) : targets[a.pair];
                bool dup = false;
                if (set)
                    dup = !set->insert(cand);
                else
                    for (uint32_t i = 0; i < k && !dup; ++i) dup = row[i] == cand;
                if (!dup) {
                    row[k] = cand;
                    break;
                }
                ++retries;
            }
        }
    }
    delete set;
    if (row_ptr)
        for (uint64_t t = 0; t <= n; ++t) row_ptr[t] = (int64_t)(t * d);
    if (retries_out) *retries_out = retries;
    return PPA_OK;
}

int ppa_cpu_degree_histogram(const ppa_graph_desc *desc, const int32_t *targets, uint32_t *degree,
                             uint64_t *hist, uint64_t hist_len, uint64_t *max_degree)
{
    int rc = ppa_check_desc(desc);
    if (rc) return rc;
    if (!targets || !degree || (hist_len && !hist)) return PPA_ERR_INVALID;
    const uint64_t n = desc->n;
    const uint32_t d = desc->d;
    for (uint64_t v = 0; v < n; ++v) degree[v] = ppa::base_degree(desc->model, d, v);
    for (uint64_t s = (uint64_t)(d + 1) * d; s < n * d; ++s) {
        const int32_t v = targets[s];
        if (v < 0 || (uint64_t)v >= n) return PPA_ERR_INVALID;
        if (degree[v] == UINT32_MAX) return PPA_ERR_RANGE;  // the degree would pass 2^32 - 1
        degree[v]++;
    }
    if (hist_len) std::memset(hist, 0, hist_len * sizeof(uint64_t));
    uint64_t mx = 0;
    for (uint64_t v = 0; v < n; ++v) {
        if (degree[v] > mx) mx = degree[v];
        if (degree[v] < hist_len) hist[degree[v]]++;
    }
    if (max_degree) *max_degree = mx;
    return PPA_OK;
}

}  // extern "C"
