// check_cpu.cpp -- host-side code under AddressSanitizer and UBSan.
//
// Built by tests/sanitize/Makefile with -fsanitize=address,undefined from
// the product's serial path (csrc/ppa_cpu.cpp: ppa_cpu_generate,
// ppa_cpu_degree_histogram, ppa_check_desc, ppa_p_threshold) and the oracle
// (oracle/ppa_oracle.c: generation, Barabasi-Albert, closure, verifier,
// degrees, in-edge and undirected CSR). Every function runs on small graphs
// of every model and on edge cases (seed clique only, d = 1, d = PPA_MAX_D,
// p = 0 and 1, planted mistakes for the verifier); the product and the oracle
// must agree. Exit status 0 = all checks held and no sanitizer report.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "ppa.h"

extern "C" {
int oracle_generate(uint64_t n, uint32_t d, uint64_t T, uint64_t seed, int32_t *G, int64_t *row_ptr,
                    uint64_t *total_retries);
int oracle_generate_ba(uint64_t n, uint32_t d, uint64_t seed, int32_t *G, uint64_t *total_retries);
int oracle_degrees(uint64_t n, uint32_t d, const int32_t *G, uint64_t *deg_out, uint64_t *hist, uint64_t hist_len,
                   uint64_t *max_degree);
int oracle_degrees_ba(uint64_t n, uint32_t d, const int32_t *G, uint64_t *deg_out, uint64_t *hist,
                      uint64_t hist_len, uint64_t *max_degree);
int oracle_in_csr(uint64_t n, uint32_t d, const int32_t *G, int64_t *in_ptr, int32_t *in_src);
int oracle_undirected_csr(uint64_t n, uint32_t d, const int32_t *G, int64_t *ptr, int32_t *adj);
int oracle_verify_parallel_model(uint32_t model, uint64_t n, uint32_t d, uint64_t T, uint64_t seed, const int32_t *G,
                                 uint64_t first, uint64_t count, int threads, uint64_t *bad);
typedef struct oracle_closure oracle_closure;
oracle_closure *oracle_closure_create(uint64_t n, uint32_t d, uint64_t T, uint64_t seed);
int oracle_closure_set_prefix(oracle_closure *c, const int32_t *prefix, uint64_t prefix_nodes);
int oracle_closure_rows(oracle_closure *c, uint64_t first_node, uint64_t count, int32_t *out);
void oracle_closure_destroy(oracle_closure *c);
}

static int failures = 0;

#define CHECK(cond)                                                          \
    do {                                                                     \
        if (!(cond)) {                                                       \
            std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #cond); \
            ++failures;                                                      \
        }                                                                    \
    } while (0)

static void one_graph(uint64_t n, uint32_t d, double p, uint64_t seed, uint32_t model)
{
    const uint64_t T = ppa_p_threshold(p);
    const uint64_t slots = n * d;
    std::vector<int32_t> want(slots), got(slots);
    std::vector<int64_t> rp(n + 1);
    uint64_t r_or = 0, r_cpu = 0;
    if (model == PPA_MODEL_COPY)
        CHECK(oracle_generate(n, d, T, seed, want.data(), nullptr, &r_or) == 0);
    else
        CHECK(oracle_generate_ba(n, d, seed, want.data(), &r_or) == 0);
    ppa_graph_desc g{};
    g.n = n;
    g.d = d;
    g.model = model;
    g.p_threshold = T;
    g.seed = seed;
    CHECK(ppa_cpu_generate(&g, got.data(), rp.data(), &r_cpu) == PPA_OK);
    CHECK(got == want);
    for (uint64_t t = 0; t <= n; ++t) CHECK(rp[t] == (int64_t)(t * d));
    // every row passes the verifier, and a planted mistake is caught
    uint64_t bad = 0;
    CHECK(oracle_verify_parallel_model(model, n, d, T, seed, got.data(), 0, n, 3, &bad) == 0);
    CHECK(bad == UINT64_MAX);
    if (n > d + 2) {
        std::vector<int32_t> planted(got);
        const uint64_t t = n - 1;
        planted[t * d] = planted[t * d] == 0 ? 1 : 0;
        CHECK(oracle_verify_parallel_model(model, n, d, T, seed, planted.data(), 0, n, 2, &bad) == 0);
        CHECK(bad == t);
    }
    // degrees: product against oracle
    std::vector<uint32_t> deg(n);
    std::vector<uint64_t> hist(4096), ohist(4096), odeg(n);
    uint64_t mx = 0, omx = 0;
    CHECK(ppa_cpu_degree_histogram(&g, got.data(), deg.data(), hist.data(), hist.size(), &mx) == PPA_OK);
    if (model == PPA_MODEL_COPY)
        CHECK(oracle_degrees(n, d, want.data(), odeg.data(), ohist.data(), ohist.size(), &omx) == 0);
    else
        CHECK(oracle_degrees_ba(n, d, want.data(), odeg.data(), ohist.data(), ohist.size(), &omx) == 0);
    CHECK(mx == omx);
    CHECK(hist == ohist);
    for (uint64_t v = 0; v < n; ++v) CHECK(deg[v] == odeg[v]);
    // in-edge CSR, and the undirected lists of a copy-model graph
    std::vector<int64_t> in_ptr(n + 1);
    std::vector<int32_t> in_src(slots);
    CHECK(oracle_in_csr(n, d, want.data(), in_ptr.data(), in_src.data()) == 0);
    CHECK(in_ptr[n] == (int64_t)slots);
    if (model == PPA_MODEL_COPY) {
        std::vector<int64_t> uptr(n + 1);
        std::vector<int32_t> adj(2 * slots);
        CHECK(oracle_undirected_csr(n, d, want.data(), uptr.data(), adj.data()) == 0);
        CHECK(uptr[n] == (int64_t)(2 * slots - (uint64_t)(d + 1) * d));
        // closure rows of the last nodes, with and without a prefix
        oracle_closure *c = oracle_closure_create(n, d, T, seed);
        CHECK(c != nullptr);
        if (c) {
            const uint64_t w = n < 64 ? n : 64;
            std::vector<int32_t> rows(w * d);
            CHECK(oracle_closure_rows(c, n - w, w, rows.data()) == 0);
            for (uint64_t i = 0; i < w * d; ++i) CHECK(rows[i] == want[(n - w) * d + i]);
            CHECK(oracle_closure_set_prefix(c, want.data(), n / 2) == 0);
            CHECK(oracle_closure_rows(c, n - w, w, rows.data()) == 0);
            for (uint64_t i = 0; i < w * d; ++i) CHECK(rows[i] == want[(n - w) * d + i]);
            oracle_closure_destroy(c);
        }
    }
}

int main()
{
    // descriptor checks and thresholds
    CHECK(ppa_p_threshold(0.0) == 0 && ppa_p_threshold(-1.0) == 0);
    CHECK(ppa_p_threshold(1.0) == (1ull << 32) && ppa_p_threshold(0.5) == (1ull << 31));
    CHECK(ppa_check_desc(nullptr) == PPA_ERR_INVALID);
    ppa_graph_desc bad{};
    bad.n = 10;
    bad.d = 0;
    CHECK(ppa_check_desc(&bad) == PPA_ERR_INVALID);
    bad.d = PPA_MAX_D + 1;
    CHECK(ppa_check_desc(&bad) == PPA_ERR_INVALID);
    bad.d = 10;
    CHECK(ppa_check_desc(&bad) == PPA_ERR_INVALID);  // n < d + 1
    bad.d = 4;
    bad.p_threshold = (1ull << 32) + 1;
    CHECK(ppa_check_desc(&bad) == PPA_ERR_INVALID);
    bad.p_threshold = 0;
    bad.model = 7;
    CHECK(ppa_check_desc(&bad) == PPA_ERR_INVALID);
    CHECK(ppa_strerror(-9) != nullptr && ppa_strerror(12345) != nullptr);

    const struct {
        uint64_t n;
        uint32_t d;
        double p;
        uint64_t seed;
    } cases[] = {
        {2, 1, 0.5, 1},       {3, 2, 0.5, 2},          {17, 16, 0.3, 3},       {33, 32, 0.1, 4},
        {2000, 1, 0.0, 5},    {2000, 1, 1.0, 6},       {5000, 3, 0.0, 7},      {20000, 8, 0.2, 8},
        {6000, 16, 0.1, 9},   {3000, 32, 0.5, 10},     {10000, 5, 0.9, 1ull << 40},
        {34, 33, 0.5, 11},    {2000, 33, 0.2, 12},     {1500, 100, 0.0, 13},   {1100, 1000, 0.7, 14},
    };
    for (const auto &c : cases) {
        one_graph(c.n, c.d, c.p, c.seed, PPA_MODEL_COPY);
        if (c.n > c.d) one_graph(c.n, c.d, 0.5, c.seed, PPA_MODEL_BA);
    }
    if (failures) {
        std::fprintf(stderr, "%d checks failed\n", failures);
        return 1;
    }
    std::printf("sanitized host checks passed\n");
    return 0;
}
