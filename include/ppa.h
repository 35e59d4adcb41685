/*
 * ppa.h -- C ABI of the MI355X copy-model preferential-attachment generator.
 *
 * This is the drop-in boundary for cuPPA's GPU generator. The reference has no
 * FFI; its path sits behind the C++ class cuPPA (src/PA/cuPPA/cuPPA.cuh:84-105)
 * whose generateGraph / generateGraphAdaptive (cuPPA_1.cu:9-254,
 * cuPPA_2.cu:9-241) allocate the graph, run the stage loop, optionally write
 * the degree distribution (cuPPA.cu:103-165) and free everything. The entry
 * points below replace that class one for one, with these differences:
 *   - the caller owns the output buffers (device pointers) and gets the graph
 *     back instead of having it freed inside the call;
 *   - no exceptions or exit() cross the ABI: every function returns an int
 *     status (PPA_OK = 0, negative = error, see ppa_strerror);
 *   - the device is never reset (cuPPA_1.cu:11 calls cudaDeviceReset);
 *   - randomness is keyed (seed, node, slot, retry) Philox-4x32-10, so the
 *     output does not depend on launch shape or scheduling.
 *
 * Model (the copy model of src/PA/SPA/SerialCopyModel3.cpp:38-78, keyed):
 *   nodes 0..d form a clique (node i's slot j = j + (j >= i)); every node
 *   t = d+1 .. n-1 fills slots k = 0..d-1 in order; attempt r of slot k draws
 *   x = Philox4x32_10(key = (lo32 seed, hi32 seed), ctr = (lo32 t, hi32 t, k, r)),
 *   direct iff x0 < p_threshold, R = mulhi64((x1 << 32) | x2, t*d),
 *   u = R / d, j = R % d, candidate = direct ? u : target[u*d + j],
 *   and the attempt is rejected (r + 1) iff the candidate already fills one of
 *   slots 0..k-1 of node t.
 *
 * Threading: a context is bound to one device. Calls on one context must not
 * overlap; different contexts may be used from different host threads.
 * Streams are passed as `void*` holding a hipStream_t (NULL = default stream).
 */
#ifndef PPA_H_
#define PPA_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------- */
enum {
    PPA_OK = 0,
    PPA_ERR_INVALID = -1,   /* bad argument (sizes, d, pointers) */
    PPA_ERR_RETRY = -2,     /* a slot exceeded PPA_MAX_RETRY rejected attempts */
    PPA_ERR_NOMEM = -3,     /* device or host allocation failed */
    PPA_ERR_DEVICE = -4,    /* HIP runtime error */
    PPA_ERR_TIMEOUT = -5,   /* the kernel watchdog expired; the output is incomplete */
    PPA_ERR_BUDGET = -6,    /* request exceeds the configured memory budget */
    PPA_ERR_NOT_FOUND = -7, /* unknown request id */
    PPA_ERR_CANCELLED = -8, /* request was cancelled */
    PPA_ERR_AGAIN = -9,     /* not ready yet / queue full, try again */
    PPA_ERR_INTERNAL = -10,
    PPA_ERR_RANGE = -11     /* a result exceeds its output type (a degree above 2^32 - 1) */
};

#define PPA_MAX_RETRY (1u << 24)
#define PPA_MAX_D 8192u          /* largest out-degree the generator supports; d <= 32 runs
                                    the lane-per-node generator, larger d one wavefront per
                                    node (the reference's 48 KB wait queue caps d at 6144,
                                    cuPPA_RRP_1.cu:49-56) */
#define PPA_MAX_BATCH_D 16u      /* largest out-degree of a batch's graphs and of requests the
                                    server may batch (larger ones always launch alone) */
#define PPA_MAX_N 0x7fffffffull  /* targets are int32 (DataTypes.h:11-18) */

/* ---- graph descriptor ---------------------------------------------------- */
/* Models (ppa_graph_desc.model, ppa_request.model):
 *  PPA_MODEL_COPY  the copy model of `-A cuppa` / `-A scm`: seed clique on
 *                  nodes 0..d, node t >= d+1 draws a uniform earlier (node,
 *                  slot) pair and takes the node with probability p, else the
 *                  slot's target (SerialCopyModel3.cpp:38-78).
 *  PPA_MODEL_BA    the Barabasi-Albert model of `-A sba` (SerialBA.cpp:68-187):
 *                  node t's d distinct targets are uniform draws over the 2*d*t
 *                  endpoints of the edges of nodes 0..t-1, where seed nodes
 *                  i < d own 2d endpoints each and node d's targets are
 *                  0..d-1. A uniform endpoint is the source or the target of a
 *                  uniform (node, slot) pair with probability 1/2 each, so
 *                  this is the copy model at p = 1/2 over seed rows i < d that
 *                  hold i (self-loops, as the reference stores them); row d
 *                  holds 0..d-1 and generation starts at d+1 as in the copy
 *                  model. p_threshold is ignored (fixed at 2^31). */
enum { PPA_MODEL_COPY = 0, PPA_MODEL_BA = 1 };

/* One graph. n >= d + 1, 1 <= d <= PPA_MAX_D, p_threshold <= 2^32, model a
 * PPA_MODEL_*. targets_offset / row_ptr_offset are element offsets into the
 * concatenated output buffers; only ppa_generate_batch reads them. */
typedef struct ppa_graph_desc {
    uint64_t n;
    uint32_t d;
    uint32_t model;         /* PPA_MODEL_* (0 = copy model) */
    uint64_t p_threshold;   /* min(2^32, floor(p * 2^32)); see ppa_p_threshold */
    uint64_t seed;          /* Philox key */
    uint64_t targets_offset;
    uint64_t row_ptr_offset;
} ppa_graph_desc;

/* Launch shape of the generator. Any shape gives the same output. For d > 32
 * the generator runs one wavefront per node in 64-thread workgroups: there
 * block_threads, chunk_nodes, early_nodes and prefix_nodes are ignored, and
 * blocks_per_cu, max_blocks, watchdog_ms and sleep_cap apply as below. */
typedef struct ppa_tuning {
    uint32_t block_threads;  /* workgroup size, multiple of 64, <= 256 (0 = default 256) */
    uint32_t blocks_per_cu;  /* resident workgroups per CU, at most the occupancy maximum (0 = that
                                maximum, and for ppa_generate* at most 16 wavefronts per CU) */
    uint32_t chunk_nodes;    /* nodes a wavefront claims at once, multiple of 64 (0 = default) */
    uint32_t max_blocks;     /* cap on the persistent grid (0 = no cap) */
    uint32_t watchdog_ms;    /* kernel deadline (0 = default 60000) */
    uint32_t early_nodes;    /* nodes below this index publish each slot as it is accepted and
                                are claimed one per lane (0 = default 2^22, 1 = none); batches
                                apply at most 16 chunks of it per graph */
    uint32_t sleep_cap;      /* longest back-off of a wavefront whose lanes all wait, in units of
                                256 clocks, 1..64 (0 = by grid size: wavefronts / 256, within
                                1..32) */
    uint32_t prefix_nodes;   /* LDS-resident prefix: a graph's first nodes generated by one
                                workgroup with the graph in LDS, where a dependent step costs an
                                LDS read, not a memory round trip. 0 = default (none when
                                d(1-p) <= 4; 40000 / d nodes for a single graph with d >= 9
                                and d(1-p) >= 8; else 1024 + d + 1), 1 = none, N = at most N
                                nodes (never more than 40000 / d for a single graph with
                                d >= 9, 16320 / d otherwise and in batches) */
} ppa_tuning;

/* Counters of the last generate call (cuPPA_1.cu:84-87 timing CSV and the
 * max-queue statistic of cuPPA_RRP_1.cu:191-208 become these). */
typedef struct ppa_stats {
    uint64_t nodes;          /* nodes generated (t >= d+1) */
    uint64_t slots;          /* slots filled */
    uint64_t draws;          /* Philox calls */
    uint64_t retries;        /* rejected attempts (duplicate targets) */
    uint64_t prefetch_misses;/* copy reads that found the source slot still empty */
    uint64_t polls;          /* coherent re-reads of an empty source slot */
    uint64_t stalls;         /* polls that found the slot still empty */
    uint64_t claims;         /* work-claim atomics */
    uint64_t sleeps;         /* wavefront back-off sleeps */
    uint64_t staged_rows;    /* rows published through the coalesced LDS path */
    uint64_t passes;         /* loop passes, summed over wavefronts (nodes / passes = lanes
                                refilled per pass) */
    uint64_t grid_blocks;    /* persistent workgroups launched */
    uint64_t block_threads;
    double kernel_ms;        /* device time of the generation kernel */
    double total_ms;         /* device time of the whole call (fill + generate) */
} ppa_stats;

typedef struct ppa_ctx ppa_ctx;

/* ---- utilities ------------------------------------------------------------ */
const char *ppa_strerror(int code);
/* min(2^32, floor(p * 2^32)); p <= 0 -> 0. */
uint64_t ppa_p_threshold(double p);
/* Validates a descriptor (same checks as ppa_generate). */
int ppa_check_desc(const ppa_graph_desc *desc);

/* ---- context (replaces the cuPPA object, cuPPA.cuh:84-105) ---------------- */
/* scratch_budget_bytes caps the device work memory the context may hold for
 * the degree pass and the in-edge / undirected CSR (0 = no cap). The
 * generators themselves need only the context's fixed 128-byte scratch.
 * Work memory is kept between calls (the next call of the same size reuses
 * it) until ppa_ctx_trim or ppa_ctx_destroy. Over the cap, the degree pass
 * counts with one atomic per slot instead of its partitioned passes (same
 * result, slower), and the CSR forms return PPA_ERR_BUDGET. */
int ppa_ctx_create(ppa_ctx **out, int device, uint64_t scratch_budget_bytes);
int ppa_ctx_destroy(ppa_ctx *ctx);
/* Frees the work memory the context holds (waits for the device first).
 * PPA_ERR_AGAIN while a launch is pending. */
int ppa_ctx_trim(ppa_ctx *ctx);
/* Work memory now held and the cap it is held to (0 = none); either nullable. */
int ppa_ctx_memory(ppa_ctx *ctx, uint64_t *held_bytes, uint64_t *budget_bytes);
int ppa_ctx_set_tuning(ppa_ctx *ctx, const ppa_tuning *tuning);
int ppa_ctx_get_tuning(ppa_ctx *ctx, ppa_tuning *out);
int ppa_get_stats(ppa_ctx *ctx, ppa_stats *out);

/* ---- single graph: replaces cuPPA::generateGraph (cuPPA_1.cu:9-254) and
 *      cuPPA::generateGraphAdaptive (cuPPA_2.cu:9-241) ---------------------- */
/* Writes d_targets[0 .. n*d) (int32, row-major [n][d]) and, if d_row_ptr is
 * not NULL, d_row_ptr[0 .. n] = t*d. Enqueues on `stream`, waits for it, and
 * returns the kernel's status. */
int ppa_generate(ppa_ctx *ctx, const ppa_graph_desc *desc, int32_t *d_targets,
                 int64_t *d_row_ptr, void *stream);
/* No counterpart in the reference. For callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc, code that must
 * not allocate between or around stream captures; the call waits for its
 * stream, so it is not itself capturable): ppa_generate on the current device
 * without a context,
 * with the default launch rules. It allocates nothing: the launch's claim
 * counter, status word and counters (128 bytes) live in a static device
 * buffer of 256 bytes that the library holds on each GPU it runs on until the
 * process ends. That record is all the generator needs whatever n and d are,
 * so no valid graph exceeds it; descriptors are checked as ppa_generate
 * checks them. Each call resets the record before its kernels run. Calls that
 * overlap on one GPU share the record and race: issue them one at a time per
 * GPU. Enqueues on `stream`, waits, and returns the kernel's status; stats
 * (nullable) receives the counters (kernel_ms and total_ms are 0: no events
 * are created). */
int ppa_generate_noalloc(const ppa_graph_desc *desc, int32_t *d_targets, int64_t *d_row_ptr,
                         ppa_stats *stats, void *stream);
/* Same work without the final wait; the status is read by ppa_ctx_sync. Until
 * then the launch owns the context's scratch: another ppa_generate*,
 * ppa_batch_launch, ppa_generate_batch, ppa_autotune, ppa_degree_histogram,
 * ppa_histogram_of_degrees, ppa_in_csr or ppa_undirected_csr call on the
 * context returns PPA_ERR_AGAIN. ppa_ctx_sync waits for the launch's own end
 * (an event recorded on the launch's stream), whichever stream it is given,
 * and frees the context whatever it returns. */
int ppa_generate_async(ppa_ctx *ctx, const ppa_graph_desc *desc, int32_t *d_targets,
                       int64_t *d_row_ptr, void *stream);
int ppa_ctx_sync(ppa_ctx *ctx, void *stream);
/* No counterpart in the reference (whose writers start after the whole graph,
 * PAHelper.hpp:163-177): for callers that copy a graph out while the rest is
 * still being generated. Enqueues, like ppa_generate_async, the generation of
 * nodes [first, last) only; ppa_ctx_sync ends it. A graph is covered by
 * consecutive calls: the first has first = 0 (it also fills all of d_targets
 * and d_row_ptr), each later one starts where the previous one ended, with
 * the same desc and d_targets, on the same stream. When ppa_ctx_sync returns,
 * rows [0, last) are final. The ranges give the graph ppa_generate gives.
 * PPA_ERR_INVALID for last <= first, last > n, or a range that does not
 * continue the previous one. A range whose ppa_ctx_sync fails cannot be
 * continued: the graph starts again at first = 0. */
int ppa_generate_range(ppa_ctx *ctx, const ppa_graph_desc *desc, int32_t *d_targets,
                       int64_t *d_row_ptr, uint64_t first, uint64_t last, void *stream);

/* ---- degree pass: replaces cuPPA::printDegreeDistribution (cuPPA.cu:103-165)
 *      and PAHelper::outputSCMDegreeDistribution (PAHelper.hpp:15-67) ------- */
/* deg(v) = d + #{slots of nodes t >= d+1 naming v} for the copy model; for
 * PPA_MODEL_BA, deg(v) = d*[v >= d] + #{slots of nodes t >= d naming v}, the
 * count of PAHelper::outputSBADegreeDistribution (PAHelper.hpp:69-147), which
 * leaves out the seed self-loops. d_degree (nullable,
 * uint32[n]) receives deg; d_hist (nullable, uint64[hist_len]) receives
 * hist[x] = #{v : deg(v) == x} for x < hist_len; d_max_degree (nullable,
 * one uint64) receives max deg. Waits for the stream. PPA_ERR_RANGE if a
 * degree exceeds 2^32 - 1 (the reference counts in int64): only a graph of
 * more than 2^32 slots can, when over 2^32 - 1 - d of them name one node;
 * the outputs then hold that degree modulo 2^32. */
int ppa_degree_histogram(ppa_ctx *ctx, const ppa_graph_desc *desc, const int32_t *d_targets,
                         uint32_t *d_degree, uint64_t *d_hist, uint64_t hist_len,
                         uint64_t *d_max_degree, void *stream);
/* The histogram step alone (cuPPA.cu:147-158), over degrees a previous
 * ppa_degree_histogram call left in d_degree (uint32[n]): d_hist[x] =
 * #{v : deg(v) == x} for x < hist_len (nullable with hist_len 0), and the
 * maximum degree. Lets a caller size the histogram from the maximum without
 * counting the slots twice. Waits for the stream. */
int ppa_histogram_of_degrees(ppa_ctx *ctx, uint64_t n, const uint32_t *d_degree, uint64_t *d_hist,
                             uint64_t hist_len, uint64_t *d_max_degree, void *stream);

/* ---- in-edge CSR (no counterpart in the reference, which never builds one;
 *      the consumer of `targets` that the reference's users build next) ---- */
/* in_src[in_ptr[v] .. in_ptr[v+1]) = the sources t of every slot (seed rows
 * included) naming v, in ascending order; d_in_ptr int64[n+1], d_in_src
 * int32[n*d]; in_ptr[n] = slots holding a node id in [0, n). A stable radix
 * sort on the device: needs 16*n*d bytes of work memory, which the context
 * keeps for the next call (PPA_ERR_NOMEM if it cannot be had). Waits for the
 * stream. */
int ppa_in_csr(ppa_ctx *ctx, const ppa_graph_desc *desc, const int32_t *d_targets, int64_t *d_in_ptr,
               int32_t *d_in_src, void *stream);
/* Undirected lists of a complete graph: d_adj[d_ptr[v] .. d_ptr[v+1]) = v's d
 * targets (slot order), then the sources t >= d+1 of the slots naming v
 * (ascending), so each seed-clique edge appears once per endpoint and
 * d_ptr[v+1] - d_ptr[v] = deg(v) of ppa_degree_histogram. d_ptr int64[n+1],
 * d_adj int32[2*n*d - (d+1)*d]. Work memory: that of ppa_in_csr plus
 * 8*(n+1) + 4*n*d bytes. PPA_ERR_INVALID if a slot is out of range or the
 * graph is not a copy-model graph (the lists are defined over the clique
 * seed). */
int ppa_undirected_csr(ppa_ctx *ctx, const ppa_graph_desc *desc, const int32_t *d_targets, int64_t *d_ptr,
                       int32_t *d_adj, void *stream);

/* ---- batched ensemble (no counterpart in the reference: it generates one
 *      graph per process run; this is cuPPA::generateGraph over many graphs
 *      in one persistent launch) ------------------------------------------- */
/* Graph i writes d_targets[descs[i].targets_offset .. + n_i*d_i) and, if
 * d_row_ptr is not NULL, d_row_ptr[descs[i].row_ptr_offset .. + n_i + 1)
 * with values targets_offset_i + t*d_i (offsets into d_targets, so each
 * graph's CSR indexes the concatenated array). Regions must not overlap, and
 * every d_i <= PPA_MAX_BATCH_D (else PPA_ERR_INVALID). Every graph's output
 * equals ppa_generate of that graph alone. */
typedef struct ppa_batch ppa_batch;
/* Uploads the descriptor table and captures reset + fill + generate into a
 * hipGraph. The batch borrows the context: one launch in flight per context. */
int ppa_batch_create(ppa_ctx *ctx, const ppa_graph_desc *descs, uint64_t count, int32_t *d_targets,
                     int64_t *d_row_ptr, ppa_batch **out);
/* Replays the graph, does not wait. PPA_ERR_AGAIN while the context has a
 * launch pending (ppa_generate_async, ppa_generate_range or another batch). */
int ppa_batch_launch(ppa_batch *batch, void *stream);
/* Waits and returns the kernel status. While it waits it watches the graphs'
 * completion flags (raised by the kernels once a graph is complete and
 * published, before the rest of the launch ends). */
int ppa_batch_sync(ppa_batch *batch, void *stream);
/* After ppa_batch_sync: done_ms[i] = host milliseconds from ppa_batch_launch
 * to the first look that found graph i complete. Graphs not seen before the
 * launch ended (a graph of the seed clique alone has nothing to generate) get
 * its end if it succeeded; after a failed launch they stay -1. count must
 * equal the batch's graph count. No counterpart in the reference: the
 * per-graph latency of a batched ensemble. */
int ppa_batch_completion_ms(ppa_batch *batch, double *done_ms, uint64_t count);
int ppa_batch_destroy(ppa_batch *batch);
/* create + launch + sync + destroy */
int ppa_generate_batch(ppa_ctx *ctx, const ppa_graph_desc *descs, uint64_t count, int32_t *d_targets,
                       int64_t *d_row_ptr, void *stream);

/* ---- autotuner (replaces the online queue-pressure doubling of
 *      cuPPA_2.cu:181-198) --------------------------------------------------- */
/* Times candidate launch shapes on the given graphs (count == 1: the single
 * graph generator; count > 1: the batch generator), sets the best on the
 * context and returns it. Results are cached in the context by workload
 * signature (n bucket, d, p bucket, batch size), so a repeated call with the
 * same signature does not re-measure. Output buffers are overwritten. */
int ppa_autotune(ppa_ctx *ctx, const ppa_graph_desc *descs, uint64_t count, int32_t *d_targets,
                 int64_t *d_row_ptr, void *stream, ppa_tuning *best);

/* ---- serving loop (no counterpart in the reference) ------------------------
 * Requests are queued, admitted first-in first-out against a device memory
 * budget (each request's targets + row_ptr live in one pool allocated at
 * creation), batched as they are admitted, and generated on up to `streams`
 * concurrent persistent launches, so requests that arrive while others run
 * overlap with them on the device. Results stay on the device until released.
 * A request's output equals ppa_generate of the same graph alone; its row_ptr
 * holds t*d (offsets into its own targets). A request batched with others is
 * reported DONE as soon as its own graph is complete, before the rest of its
 * launch. Calls must not overlap (one host thread drives a server). */
typedef struct ppa_server ppa_server;

typedef struct ppa_server_config {
    uint64_t budget_bytes; /* device pool for request outputs */
    uint32_t streams;      /* concurrent launches (1..16, 0 = 16); they overlap only as far as the
                              HIP runtime has hardware queues (GPU_MAX_HW_QUEUES, default 4) */
    uint32_t max_batch;    /* requests per launch (0 = 4096) */
    uint64_t batch_nodes;  /* a launch closes once it holds this many nodes (0 = 2^22) */
    uint64_t solo_nodes;   /* requests of at least this many nodes launch alone (0 = 10000) */
    uint32_t plain;        /* 1: one request per launch, one stream, single-graph kernel */
    uint32_t solo_grid_div; /* a launch of one request gets 1/solo_grid_div of the workgroups a
                               single ppa_generate call would (0 = 3; 1 = the same): its
                               chain-bound workgroups hold wave slots that concurrent launches
                               could use; ignored in plain mode */
    uint32_t keep_results;  /* 1: ppa_server_replay leaves finished requests in place for the
                               caller to poll and release (their memory stays held, so the
                               budget must cover every request of the trace); 0: it releases
                               each request as soon as it finishes */
    uint32_t reserved;
} ppa_server_config;

typedef struct ppa_request {
    uint64_t id;           /* caller-chosen, unique among live requests */
    uint64_t n;
    uint32_t d;
    uint32_t model;        /* PPA_MODEL_* */
    uint64_t p_threshold;
    uint64_t seed;
} ppa_request;

enum { PPA_REQ_QUEUED = 0, PPA_REQ_RUNNING = 1, PPA_REQ_DONE = 2, PPA_REQ_FAILED = 3, PPA_REQ_CANCELLED = 4 };

typedef struct ppa_result {
    int32_t state;         /* PPA_REQ_* */
    int32_t status;        /* PPA_OK or the launch's error code */
    int32_t *targets;      /* device, n*d, valid while DONE and not released */
    int64_t *row_ptr;      /* device, n+1 */
    uint64_t n;
    uint32_t d;
    uint32_t reserved;
    double submit_ms;      /* host steady clock, ms since the server was created */
    double start_ms;       /* launch enqueued */
    double done_ms;        /* completion observed by ppa_server_step */
} ppa_result;

int ppa_server_create(ppa_ctx *ctx, const ppa_server_config *cfg, ppa_server **out);
int ppa_server_destroy(ppa_server *srv);
/* PPA_ERR_BUDGET if the request can never fit the budget; PPA_ERR_INVALID on a
 * bad descriptor or a duplicate id. */
int ppa_server_submit(ppa_server *srv, const ppa_request *req);
/* Non-blocking progress: retires finished launches, admits queued requests
 * and starts launches on free streams. Returns the number of requests that
 * finished in this call (>= 0) or an error code. */
int ppa_server_step(ppa_server *srv);
int ppa_server_poll(ppa_server *srv, uint64_t id, ppa_result *out);
int ppa_server_release(ppa_server *srv, uint64_t id);  /* frees a finished request's memory */
int ppa_server_cancel(ppa_server *srv, uint64_t id);   /* only while still queued */
int ppa_server_usage(ppa_server *srv, uint64_t *used_bytes, uint64_t *budget_bytes, uint64_t *queued,
                     uint64_t *running);
/* Replays a trace in real time: request i is submitted when the clock reaches
 * arrival_ms[i] (ms from the start of the call) and released as soon as it
 * finishes (kept for the caller with keep_results; then PPA_ERR_BUDGET if the
 * kept requests leave no room for the next one); latency_ms[i] = finish -
 * arrival, wait_ms[i] (nullable) = launch start - arrival. makespan_ms = last
 * finish. */
int ppa_server_replay(ppa_server *srv, const ppa_request *reqs, const double *arrival_ms, uint64_t count,
                      double *latency_ms, double *wait_ms, double *makespan_ms);

/* ---- serial CPU path (`-A scm`): replaces SerialCopyModel::generateGraph
 *      (src/PA/SPA/SerialCopyModel3.cpp:11-107) ---------------------------- */
/* Host memory. Same models and draws as ppa_generate, hence the same graph.
 * retries (nullable) receives the number of rejected attempts. */
int ppa_cpu_generate(const ppa_graph_desc *desc, int32_t *targets, int64_t *row_ptr, uint64_t *retries);
/* Host-memory degree pass of PAHelper::outputSCMDegreeDistribution
 * (PAHelper.hpp:15-67) with deg(v) as in ppa_degree_histogram, and
 * PPA_ERR_RANGE as there. degree is uint32[n] scratch supplied by the
 * caller. */
int ppa_cpu_degree_histogram(const ppa_graph_desc *desc, const int32_t *targets, uint32_t *degree,
                             uint64_t *hist, uint64_t hist_len, uint64_t *max_degree);

#ifdef __cplusplus
}
#endif

#endif /* PPA_H_ */
