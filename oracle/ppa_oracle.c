/*
 * CPU ORACLE -- test infrastructure only.
 *
 * This file is the checker for the copy-model generator. Only tests/, the
 * smoke() check in __graft_entry__.py and the cpu_baseline leg of bench.py
 * may load it, and only to compute expected results. It is never part of the
 * product path.
 *
 * It restates, sequentially and in plain C, the serial copy model of
 * src/PA/SPA/SerialCopyModel3.cpp:38-78 (loop over nodes, redraw until the
 * target is unique among the node's filled slots, copy = G[R]) with the
 * model's randomness replaced by keyed Philox-4x32-10 draws:
 *
 *   key     = (lo32(seed), hi32(seed))
 *   counter = (lo32(t), hi32(t), k, r)           one call per attempt
 *   direct  <=> x0 < T,   T = min(2^32, floor(p * 2^32))
 *   R       = mulhi64((x1 << 32) | x2, t * d)   a uniform (node, slot) pair
 *   u = R / d, j = R % d                        SerialCopyModel3.cpp:62-64
 *   cand    = direct ? u : G[u * d + j]          SerialCopyModel3.cpp:67-74
 *   redraw (r + 1) while cand is among G[t*d .. t*d+k-1]
 *                                               SerialCopyModel3.cpp:60-75,
 *                                               BSTVertex.h:189-235
 *
 * The seed graph is the clique on nodes 0..d: node i's slot j holds
 * j + (j >= i). Generation starts at t = d + 1. Slots are initialised to -1
 * first (SerialCopyModel.cpp:30-44). Degree follows cuPPA.cu:115-131 and
 * PAHelper.hpp:25-45: deg(v) = d (own slots) + number of slots that name v.
 *
 * Parity pins (see tests/): Philox against the published Random123
 * known-answer vectors and an independent Python big-integer restatement;
 * the mappings against Python big integers; generation against closed-form
 * special cases (p = 1, p = 0, d = 1), invariants and a pure-Python
 * restatement on small inputs; the closure oracle against the sequential one;
 * the in-edge CSR against an independent stable argsort and against the
 * degree pass.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORACLE_OK 0
#define ORACLE_ERR_ARG -1
#define ORACLE_ERR_RETRY -2
#define ORACLE_ERR_NOMEM -3

/* Retry cap per slot; the product applies the same cap. */
#define ORACLE_MAX_RETRY (1u << 24)

/* Largest out-degree the row verifier takes (the product's limit too). */
#define ORACLE_MAX_D 8192u

/* ---- Philox-4x32-10 (Salmon et al., SC'11) ------------------------------ */

static const uint32_t PHILOX_M0 = 0xD2511F53u;
static const uint32_t PHILOX_M1 = 0xCD9E8D57u;
static const uint32_t PHILOX_W0 = 0x9E3779B9u;
static const uint32_t PHILOX_W1 = 0xBB67AE85u;

void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += PHILOX_W0;
        k1 += PHILOX_W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* ---- mappings ------------------------------------------------------------ */

uint64_t oracle_p_threshold(double p)
{
    if (!(p > 0.0)) return 0;
    if (p >= 1.0) return 1ull << 32;
    /* p * 2^32 is exact in double (scaling by a power of two) */
    return (uint64_t)(p * 4294967296.0);
}

uint64_t oracle_mulhi64(uint64_t a, uint64_t b)
{
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
}

typedef struct {
    int direct;
    uint64_t u;
    uint32_t j;
} oracle_draw;

/* One attempt for slot (t, k), retry r. */
void oracle_attempt(uint64_t seed, uint64_t T, uint64_t t, uint32_t d, uint32_t k, uint32_t r,
                    int *direct, uint64_t *u, uint32_t *j)
{
    uint32_t ctr[4] = {(uint32_t)t, (uint32_t)(t >> 32), k, r};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t x[4];
    oracle_philox4x32_10(ctr, key, x);
    uint64_t R = oracle_mulhi64(((uint64_t)x[1] << 32) | x[2], t * (uint64_t)d);
    *direct = (uint64_t)x[0] < T;
    *u = R / d;
    *j = (uint32_t)(R % d);
}

/* ---- sequential generator (SerialCopyModel3.cpp:11-107) ------------------ */

static void seed_clique(uint32_t d, int32_t *G)
{
    for (uint32_t i = 0; i <= d; ++i)
        for (uint32_t j = 0; j < d; ++j)
            G[(uint64_t)i * d + j] = (int32_t)(j + (j >= i));
}

/*
 * Fills targets[0 .. n*d) (row-major [n][d]) and, if row_ptr is non-NULL,
 * row_ptr[0 .. n] = t*d. total_retries (nullable) receives the number of
 * rejected attempts over the whole graph.
 */
int oracle_generate(uint64_t n, uint32_t d, uint64_t T, uint64_t seed, int32_t *G,
                    int64_t *row_ptr, uint64_t *total_retries)
{
    if (d == 0 || n < (uint64_t)d + 1 || n > 0x7fffffffull || T > (1ull << 32) || !G)
        return ORACLE_ERR_ARG;
    uint64_t slots = n * d;
    for (uint64_t s = 0; s < slots; ++s) G[s] = -1;
    seed_clique(d, G);
    uint64_t retries = 0;
    for (uint64_t t = (uint64_t)d + 1; t < n; ++t) {
        int32_t *row = G + t * d;
        for (uint32_t k = 0; k < d; ++k) {
            uint32_t r = 0;
            for (;;) {
                int direct;
                uint64_t u;
                uint32_t j;
                oracle_attempt(seed, T, t, d, k, r, &direct, &u, &j);
                int32_t cand = direct ? (int32_t)u : G[u * d + j];
                int dup = 0;
                for (uint32_t i = 0; i < k; ++i)
                    if (row[i] == cand) { dup = 1; break; }
                if (!dup) { row[k] = cand; break; }
                if (++r >= ORACLE_MAX_RETRY) return ORACLE_ERR_RETRY;
                ++retries;
            }
        }
    }
    if (row_ptr)
        for (uint64_t t = 0; t <= n; ++t) row_ptr[t] = (int64_t)(t * d);
    if (total_retries) *total_retries = retries;
    return ORACLE_OK;
}

/* ---- serial Barabasi-Albert (SerialBA.cpp:68-187) ------------------------ */
/*
 * The reference keeps every edge as an endpoint pair in E[2*M*N]
 * (SerialBA.cpp:26-29): seed nodes i < M hold 2M copies of i
 * (SerialBA.cpp:90-97); node t >= M holds (target_j, t) at E[2Mt + 2j],
 * E[2Mt + 2j + 1] (SerialBA.cpp:138-145). Node M's targets are 0..M-1 in
 * order (the BSTVertex seeded at SerialBA.cpp:108-112). Node t+1's targets
 * are distinct draws of E[UniformMax(Z)] over the Z = 2M(t+1) endpoints
 * written so far, kept in draw order (SerialBA.cpp:149-155, BSTVertex.h:
 * 189-235). Here the uniform endpoint index is e = 2R + b from the keyed draw
 * of attempt r of slot k of node t: R = mulhi64((x1 << 32) | x2, t*M) is a
 * uniform index below t*M and b = (x0 < 2^31) a fair bit, so e is uniform
 * below 2Mt. The same retry cap as the copy model applies.
 *
 * Output is the targets of every node in the [n][d] layout of the copy model
 * (E's even entries): rows i < d hold i, row d holds 0..d-1.
 */
int oracle_generate_ba(uint64_t n, uint32_t d, uint64_t seed, int32_t *G, uint64_t *total_retries)
{
    if (d == 0 || n < (uint64_t)d + 1 || n > 0x7fffffffull || !G) return ORACLE_ERR_ARG;
    const uint64_t M = d;
    int32_t *E = (int32_t *)malloc(2 * n * M * sizeof(int32_t));
    if (!E) return ORACLE_ERR_NOMEM;
    for (uint64_t i = 0; i < M; ++i)
        for (uint64_t x = 0; x < 2 * M; ++x) E[2 * M * i + x] = (int32_t)i;
    for (uint64_t j = 0; j < M; ++j) {
        E[2 * M * M + 2 * j] = (int32_t)j;
        E[2 * M * M + 2 * j + 1] = (int32_t)M;
    }
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint64_t retries = 0;
    int rc = ORACLE_OK;
    for (uint64_t t = M + 1; t < n && rc == ORACLE_OK; ++t) {
        int32_t *pair = E + 2 * M * t;
        for (uint32_t k = 0; k < M && rc == ORACLE_OK; ++k) {
            for (uint32_t r = 0;; ++r) {
                const uint32_t ctr[4] = {(uint32_t)t, (uint32_t)(t >> 32), k, r};
                uint32_t x[4];
                oracle_philox4x32_10(ctr, key, x);
                const uint64_t R = oracle_mulhi64(((uint64_t)x[1] << 32) | x[2], t * M);
                const uint64_t e = 2 * R + (x[0] < (1u << 31) ? 1 : 0);
                const int32_t cand = E[e];
                int dup = 0;
                for (uint32_t i = 0; i < k; ++i)
                    if (pair[2 * i] == cand) { dup = 1; break; }
                if (!dup) {
                    pair[2 * k] = cand;
                    pair[2 * k + 1] = (int32_t)t;
                    break;
                }
                if (r + 1 >= ORACLE_MAX_RETRY) { rc = ORACLE_ERR_RETRY; break; }
                ++retries;
            }
        }
    }
    if (rc == ORACLE_OK)
        for (uint64_t s = 0; s < n * M; ++s) G[s] = E[2 * s];
    free(E);
    if (total_retries) *total_retries = retries;
    return rc;
}

/*
 * Degrees as PAHelper.hpp:69-147 counts them for the Barabasi-Albert graph:
 * both endpoints of every pair of nodes u >= M, so
 * deg[v] = M*[v >= M] + #{slots of nodes u >= M naming v}.
 */
int oracle_degrees_ba(uint64_t n, uint32_t d, const int32_t *G, uint64_t *deg_out, uint64_t *hist,
                      uint64_t hist_len, uint64_t *max_degree)
{
    uint64_t *deg = deg_out ? deg_out : (uint64_t *)malloc(n * sizeof(uint64_t));
    if (!deg) return ORACLE_ERR_NOMEM;
    memset(deg, 0, n * sizeof(uint64_t));
    for (uint64_t u = d; u < n; ++u)
        for (uint32_t j = 0; j < d; ++j) {
            deg[G[u * d + j]]++;
            deg[u]++;
        }
    uint64_t mx = 0;
    if (hist) memset(hist, 0, hist_len * sizeof(uint64_t));
    for (uint64_t v = 0; v < n; ++v) {
        if (deg[v] > mx) mx = deg[v];
        if (hist && deg[v] < hist_len) hist[deg[v]]++;
    }
    if (max_degree) *max_degree = mx;
    if (!deg_out) free(deg);
    return ORACLE_OK;
}

/* ---- closure oracle: target(t, k) by memoised recursion ------------------ */
/*
 * Computes single slots of a huge graph without generating the prefix. Every
 * slot is a deterministic function of the keyed draws and of the slots it
 * copies from, so target(t, k) needs only row t's slots 0..k-1 and, for each
 * copy attempt, target(u, j) with u < t. Results are memoised in an
 * open-addressing table owned by the caller-visible handle.
 */
typedef struct {
    uint64_t *keys;   /* slot index + 1 (0 = empty) */
    int32_t *vals;
    uint64_t cap;     /* power of two */
    uint64_t used;
    uint64_t n, T, seed;
    uint32_t d;
    const int32_t *prefix;   /* optional: rows [0, prefix_nodes) from the sequential oracle */
    uint64_t prefix_nodes;
} oracle_closure;

static uint64_t mix64(uint64_t x)
{
    x ^= x >> 33; x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return x;
}

oracle_closure *oracle_closure_create(uint64_t n, uint32_t d, uint64_t T, uint64_t seed)
{
    if (d == 0 || n < (uint64_t)d + 1 || n > 0x7fffffffull || T > (1ull << 32)) return NULL;
    oracle_closure *c = (oracle_closure *)calloc(1, sizeof(*c));
    if (!c) return NULL;
    c->cap = 1u << 16;
    c->keys = (uint64_t *)calloc(c->cap, sizeof(uint64_t));
    c->vals = (int32_t *)malloc(c->cap * sizeof(int32_t));
    if (!c->keys || !c->vals) { free(c->keys); free(c->vals); free(c); return NULL; }
    c->n = n; c->d = d; c->T = T; c->seed = seed;
    return c;
}

/* Rows below prefix_nodes are read from `prefix` (rows 0 .. prefix_nodes-1 of
 * the same graph, produced by oracle_generate) instead of being recursed into.
 * The array must outlive the handle. */
int oracle_closure_set_prefix(oracle_closure *c, const int32_t *prefix, uint64_t prefix_nodes)
{
    if (!c || (prefix_nodes && !prefix) || prefix_nodes > c->n) return ORACLE_ERR_ARG;
    c->prefix = prefix;
    c->prefix_nodes = prefix_nodes;
    return ORACLE_OK;
}

void oracle_closure_destroy(oracle_closure *c)
{
    if (!c) return;
    free(c->keys); free(c->vals); free(c);
}

static int closure_find(const oracle_closure *c, uint64_t s, int32_t *v)
{
    uint64_t m = c->cap - 1, h = mix64(s) & m;
    for (;;) {
        uint64_t k = c->keys[h];
        if (k == 0) return 0;
        if (k == s + 1) { *v = c->vals[h]; return 1; }
        h = (h + 1) & m;
    }
}

static int closure_put(oracle_closure *c, uint64_t s, int32_t v)
{
    if ((c->used + 1) * 2 > c->cap) {
        uint64_t ncap = c->cap * 2;
        uint64_t *nk = (uint64_t *)calloc(ncap, sizeof(uint64_t));
        int32_t *nv = (int32_t *)malloc(ncap * sizeof(int32_t));
        if (!nk || !nv) { free(nk); free(nv); return ORACLE_ERR_NOMEM; }
        for (uint64_t i = 0; i < c->cap; ++i) {
            if (!c->keys[i]) continue;
            uint64_t h = mix64(c->keys[i] - 1) & (ncap - 1);
            while (nk[h]) h = (h + 1) & (ncap - 1);
            nk[h] = c->keys[i]; nv[h] = c->vals[i];
        }
        free(c->keys); free(c->vals);
        c->keys = nk; c->vals = nv; c->cap = ncap;
    }
    uint64_t m = c->cap - 1, h = mix64(s) & m;
    while (c->keys[h]) h = (h + 1) & m;
    c->keys[h] = s + 1; c->vals[h] = v; c->used++;
    return ORACLE_OK;
}

/* Recursion depth is bounded by the copy-chain length, which is small
 * (logarithmic in t in expectation) for every configuration used. */
int oracle_closure_target(oracle_closure *c, uint64_t t, uint32_t k, int32_t *out)
{
    const uint32_t d = c->d;
    if (t >= c->n || k >= d) return ORACLE_ERR_ARG;
    if (t <= d) { *out = (int32_t)(k + (k >= t)); return ORACLE_OK; }
    uint64_t s = t * d + k;
    if (t < c->prefix_nodes) { *out = c->prefix[s]; return ORACLE_OK; }
    if (closure_find(c, s, out)) return ORACLE_OK;
    int32_t prev[64];
    int32_t *row = prev;
    int32_t *heap = NULL;
    if (k > 64) {
        heap = (int32_t *)malloc(sizeof(int32_t) * k);
        if (!heap) return ORACLE_ERR_NOMEM;
        row = heap;
    }
    int rc = ORACLE_OK;
    for (uint32_t i = 0; i < k && rc == ORACLE_OK; ++i) rc = oracle_closure_target(c, t, i, &row[i]);
    uint32_t r = 0;
    int32_t cand = -1;
    while (rc == ORACLE_OK) {
        int direct;
        uint64_t u;
        uint32_t j;
        oracle_attempt(c->seed, c->T, t, d, k, r, &direct, &u, &j);
        if (direct) cand = (int32_t)u;
        else rc = oracle_closure_target(c, u, j, &cand);
        if (rc != ORACLE_OK) break;
        int dup = 0;
        for (uint32_t i = 0; i < k; ++i)
            if (row[i] == cand) { dup = 1; break; }
        if (!dup) break;
        if (++r >= ORACLE_MAX_RETRY) rc = ORACLE_ERR_RETRY;
    }
    free(heap);
    if (rc != ORACLE_OK) return rc;
    rc = closure_put(c, s, cand);
    if (rc != ORACLE_OK) return rc;
    *out = cand;
    return ORACLE_OK;
}

/* Fills out[0 .. count*d) with rows first_node .. first_node+count-1. */
int oracle_closure_rows(oracle_closure *c, uint64_t first_node, uint64_t count, int32_t *out)
{
    for (uint64_t i = 0; i < count; ++i)
        for (uint32_t k = 0; k < c->d; ++k) {
            int rc = oracle_closure_target(c, first_node + i, k, &out[i * c->d + k]);
            if (rc != ORACLE_OK) return rc;
        }
    return ORACLE_OK;
}

uint64_t oracle_closure_size(const oracle_closure *c) { return c->used; }

/* ---- degrees (cuPPA.cu:115-131, PAHelper.hpp:25-63) ----------------------- */
/*
 * deg[v] = d + #{slots naming v}; hist[x] = #{v : deg[v] == x} for x < hist_len.
 * Returns the maximum degree through max_degree. deg may be NULL (then a
 * temporary is used).
 */
int oracle_degrees(uint64_t n, uint32_t d, const int32_t *G, uint64_t *deg_out,
                   uint64_t *hist, uint64_t hist_len, uint64_t *max_degree)
{
    uint64_t *deg = deg_out ? deg_out : (uint64_t *)malloc(n * sizeof(uint64_t));
    if (!deg) return ORACLE_ERR_NOMEM;
    for (uint64_t v = 0; v < n; ++v) deg[v] = d;
    for (uint64_t t = (uint64_t)d + 1; t < n; ++t)
        for (uint32_t k = 0; k < d; ++k) deg[G[t * d + k]]++;
    uint64_t mx = 0;
    if (hist) memset(hist, 0, hist_len * sizeof(uint64_t));
    for (uint64_t v = 0; v < n; ++v) {
        if (deg[v] > mx) mx = deg[v];
        if (hist && deg[v] < hist_len) hist[deg[v]]++;
    }
    if (max_degree) *max_degree = mx;
    if (!deg_out) free(deg);
    return ORACLE_OK;
}

/* ---- in-edge CSR ---------------------------------------------------------- */
/*
 * Counting sort of all n*d slots (seed rows included) by target, in slot
 * order: in_src[in_ptr[v] .. in_ptr[v+1]) are the sources t of the slots
 * naming v, ascending. Slots outside [0, n) are skipped. No reference
 * counterpart (the reference never builds in-edge lists).
 */
int oracle_in_csr(uint64_t n, uint32_t d, const int32_t *G, int64_t *in_ptr, int32_t *in_src)
{
    const uint64_t slots = n * (uint64_t)d;
    int64_t *cur = (int64_t *)calloc(n + 1, sizeof(int64_t));
    if (!cur) return ORACLE_ERR_NOMEM;
    for (uint64_t s = 0; s < slots; ++s)
        if (G[s] >= 0 && (uint64_t)G[s] < n) cur[G[s] + 1]++;
    for (uint64_t v = 0; v < n; ++v) cur[v + 1] += cur[v];
    for (uint64_t v = 0; v <= n; ++v) in_ptr[v] = cur[v];
    for (uint64_t s = 0; s < slots; ++s)
        if (G[s] >= 0 && (uint64_t)G[s] < n) in_src[cur[G[s]]++] = (int32_t)(s / d);
    free(cur);
    return ORACLE_OK;
}

/*
 * Undirected lists: node v's d targets (slot order), then the sources t >= d+1
 * of the slots naming v, ascending. Seed-row sources are left out, so each
 * seed-clique edge appears once per endpoint and |list(v)| = deg(v) of
 * oracle_degrees. ptr int64[n+1]; adj int32[2*n*d - (d+1)*d]. G complete.
 */
int oracle_undirected_csr(uint64_t n, uint32_t d, const int32_t *G, int64_t *ptr, int32_t *adj)
{
    int64_t *in_ptr = (int64_t *)malloc((n + 1) * sizeof(int64_t));
    int32_t *in_src = (int32_t *)malloc(n * (uint64_t)d * sizeof(int32_t));
    if (!in_ptr || !in_src) {
        free(in_ptr);
        free(in_src);
        return ORACLE_ERR_NOMEM;
    }
    int rc = oracle_in_csr(n, d, G, in_ptr, in_src);
    int64_t o = 0;
    for (uint64_t v = 0; rc == ORACLE_OK && v < n; ++v) {
        ptr[v] = o;
        for (uint32_t k = 0; k < d; ++k) adj[o++] = G[v * d + k];
        for (int64_t i = in_ptr[v]; i < in_ptr[v + 1]; ++i)
            if ((uint32_t)in_src[i] > d) adj[o++] = in_src[i];
    }
    ptr[n] = o;
    free(in_ptr);
    free(in_src);
    return rc;
}

/* ---- row-consistency verifier --------------------------------------------- */
/*
 * Checks rows [first, first + count) of a complete candidate graph G: each row
 * must equal what the model produces for node t when its copy reads are served
 * from G's own earlier rows (SerialCopyModel3.cpp:57-77 for one node). Rows
 * 0..d must be the seed clique. If every row of a graph passes, the graph
 * equals oracle_generate's output by induction over t (the first row that
 * differs from the oracle would have all of its copy sources right and so
 * could not pass). Returns the first failing row through *bad (or UINT64_MAX).
 */
/*
 * model 1 checks a Barabasi-Albert graph in the copy-model form that
 * oracle_generate_ba is pinned to: seed rows i < d hold i, row d holds
 * 0..d-1, and the coin threshold is 2^31 (T is ignored).
 */
int oracle_verify_rows_model(uint32_t model, uint64_t n, uint32_t d, uint64_t T, uint64_t seed,
                             const int32_t *G, uint64_t first, uint64_t count, uint64_t *bad)
{
    *bad = UINT64_MAX;
    if (d == 0 || d > ORACLE_MAX_D || n < (uint64_t)d + 1 || first + count > n || model > 1) return ORACLE_ERR_ARG;
    if (model == 1) T = 1ull << 31;
    int32_t row[ORACLE_MAX_D];
    for (uint64_t t = first; t < first + count; ++t) {
        const int32_t *g = G + t * d;
        if (t <= d) {
            for (uint32_t k = 0; k < d; ++k) {
                const int32_t want = model == 1 && t < d ? (int32_t)t : (int32_t)(k + (k >= t));
                if (g[k] != want) { *bad = t; return ORACLE_OK; }
            }
            continue;
        }
        for (uint32_t k = 0; k < d; ++k) {
            uint32_t r = 0;
            int32_t cand;
            for (;;) {
                int direct;
                uint64_t u;
                uint32_t j;
                oracle_attempt(seed, T, t, d, k, r, &direct, &u, &j);
                cand = direct ? (int32_t)u : G[u * d + j];
                int dup = 0;
                for (uint32_t i = 0; i < k; ++i)
                    if (row[i] == cand) { dup = 1; break; }
                if (!dup) break;
                if (++r >= ORACLE_MAX_RETRY) { *bad = t; return ORACLE_OK; }
            }
            if (cand != g[k]) { *bad = t; return ORACLE_OK; }
            row[k] = cand;
        }
    }
    return ORACLE_OK;
}

int oracle_verify_rows(uint64_t n, uint32_t d, uint64_t T, uint64_t seed, const int32_t *G,
                       uint64_t first, uint64_t count, uint64_t *bad)
{
    return oracle_verify_rows_model(0, n, d, T, seed, G, first, count, bad);
}

typedef struct {
    uint64_t n, T, seed, first, count, bad;
    uint32_t d, model;
    const int32_t *G;
    int rc;
} verify_job;

static void *verify_worker(void *arg)
{
    verify_job *j = (verify_job *)arg;
    j->rc = oracle_verify_rows_model(j->model, j->n, j->d, j->T, j->seed, j->G, j->first, j->count, &j->bad);
    return NULL;
}

int oracle_verify_parallel_model(uint32_t model, uint64_t n, uint32_t d, uint64_t T, uint64_t seed,
                                 const int32_t *G, uint64_t first, uint64_t count, int threads, uint64_t *bad);

/* oracle_verify_rows over [first, first+count) split across `threads` threads. */
int oracle_verify_parallel(uint64_t n, uint32_t d, uint64_t T, uint64_t seed, const int32_t *G,
                           uint64_t first, uint64_t count, int threads, uint64_t *bad)
{
    return oracle_verify_parallel_model(0, n, d, T, seed, G, first, count, threads, bad);
}

int oracle_verify_parallel_model(uint32_t model, uint64_t n, uint32_t d, uint64_t T, uint64_t seed,
                                 const int32_t *G, uint64_t first, uint64_t count, int threads, uint64_t *bad)
{
    *bad = UINT64_MAX;
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    verify_job jobs[256];
    pthread_t tid[256];
    uint64_t per = (count + threads - 1) / threads;
    int started = 0;
    for (int i = 0; i < threads; ++i) {
        uint64_t lo = first + (uint64_t)i * per;
        if (lo >= first + count) break;
        uint64_t hi = lo + per < first + count ? lo + per : first + count;
        jobs[i] = (verify_job){n, T, seed, lo, hi - lo, UINT64_MAX, d, model, G, 0};
        if (pthread_create(&tid[i], NULL, verify_worker, &jobs[i]) != 0) {
            verify_worker(&jobs[i]);
            tid[i] = 0;
        }
        started = i + 1;
    }
    int rc = ORACLE_OK;
    for (int i = 0; i < started; ++i) {
        if (tid[i]) pthread_join(tid[i], NULL);
        if (jobs[i].rc != ORACLE_OK) rc = jobs[i].rc;
        if (jobs[i].bad < *bad) *bad = jobs[i].bad;
    }
    return rc;
}
