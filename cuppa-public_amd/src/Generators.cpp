// Generators.cpp -- cuPPA and SerialCopyModel over the C ABI.
#include "Generators.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <cstdio>
#include <fstream>
#include <iostream>

std::string joinPath(const std::string &dir, const std::string &file)
{
    if (dir.empty()) return file;
    return dir.back() == '/' ? dir + file : dir + "/" + file;
}

bool writeDegreeCsv(const std::string &path, const std::vector<uint64_t> &hist)
{
    std::ofstream out(path);
    if (!out) return false;
    out << "Degree,Count\n";
    for (size_t deg = 0; deg < hist.size(); ++deg)
        if (hist[deg]) out << deg << "," << hist[deg] << "\n";
    return bool(out);
}

bool writeEdgeList(const std::string &path, const int32_t *G, uint64_t n, uint32_t d, uint32_t model)
{
    FILE *f = std::fopen(path.c_str(), "w");
    if (!f) return false;
    std::vector<char> buf(1 << 20);
    std::setvbuf(f, buf.data(), _IOFBF, buf.size());
    if (model == PPA_MODEL_COPY) {
        for (uint64_t i = 1; i <= d && i < n; ++i)
            for (uint32_t j = 0; j < d; ++j)
                if ((uint64_t)G[i * d + j] < i) std::fprintf(f, "%llu %d\n", (unsigned long long)i, G[i * d + j]);
    } else {
        for (uint32_t k = 0; k < d && d < n; ++k)
            std::fprintf(f, "%llu %d\n", (unsigned long long)d, G[(uint64_t)d * d + k]);
    }
    for (uint64_t t = (uint64_t)d + 1; t < n; ++t)
        for (uint32_t k = 0; k < d; ++k) std::fprintf(f, "%llu %d\n", (unsigned long long)t, G[t * d + k]);
    const bool ok = std::ferror(f) == 0;
    return std::fclose(f) == 0 && ok;
}

namespace {

// Writes device memory to a file, double-buffered through two pinned 64 MB
// buffers: the copy of one chunk runs while the previous one is written. The
// buffers are allocated once per writer (pinned allocations are slow).
class DeviceWriter {
public:
    DeviceWriter()
    {
        ok_ = hipHostMalloc(&buf_[0], kChunk, hipHostMallocDefault) == hipSuccess &&
              hipHostMalloc(&buf_[1], kChunk, hipHostMallocDefault) == hipSuccess &&
              hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&ev_[0], hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&ev_[1], hipEventDisableTiming) == hipSuccess;
    }
    ~DeviceWriter()
    {
        if (stream_) (void)hipStreamSynchronize(stream_);
        for (int i = 0; i < 2; ++i) {
            if (ev_[i]) (void)hipEventDestroy(ev_[i]);
            if (buf_[i]) (void)hipHostFree(buf_[i]);
        }
        if (stream_) (void)hipStreamDestroy(stream_);
    }
    DeviceWriter(const DeviceWriter &) = delete;
    DeviceWriter &operator=(const DeviceWriter &) = delete;

    bool write(FILE *f, const void *dptr, size_t bytes)
    {
        const char *src = (const char *)dptr;
        const size_t chunks = (bytes + kChunk - 1) / kChunk;
        auto issue = [&](size_t c) {
            const size_t off = c * kChunk, len = std::min(kChunk, bytes - off);
            return hipMemcpyAsync(buf_[c & 1], src + off, len, hipMemcpyDeviceToHost, stream_) == hipSuccess &&
                   hipEventRecord(ev_[c & 1], stream_) == hipSuccess;
        };
        bool ok = ok_;
        if (ok && chunks) ok = issue(0);
        for (size_t c = 0; ok && c < chunks; ++c) {
            ok = hipEventSynchronize(ev_[c & 1]) == hipSuccess;
            if (ok && c + 1 < chunks) ok = issue(c + 1);  // next copy runs while this chunk is written
            const size_t len = std::min(kChunk, bytes - c * kChunk);
            if (ok) ok = std::fwrite(buf_[c & 1], 1, len, f) == len;
        }
        if (stream_) (void)hipStreamSynchronize(stream_);  // no copy outlives the call
        return ok;
    }

private:
    static constexpr size_t kChunk = 64u << 20;
    void *buf_[2] = {nullptr, nullptr};
    hipStream_t stream_ = nullptr;
    hipEvent_t ev_[2] = {nullptr, nullptr};
    bool ok_ = false;
};

}  // namespace

bool writeCsr(const std::string &path, const CsrHeader &h, const int64_t *offsets, const int32_t *indices,
              bool device)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1;
    const size_t ob = (h.n + 1) * sizeof(int64_t), ib = h.nnz * sizeof(int32_t);
    if (device) {
        DeviceWriter w;
        ok = ok && w.write(f, offsets, ob) && w.write(f, indices, ib);
    } else {
        ok = ok && std::fwrite(offsets, 1, ob, f) == ob && std::fwrite(indices, 1, ib, f) == ib;
    }
    return std::fclose(f) == 0 && ok;
}

CsrHeader csrHeader(const ppa_graph_desc &g, uint64_t nnz, uint32_t kind)
{
    CsrHeader h{};
    std::memcpy(h.magic, "PPACSR1", 8);
    h.n = g.n;
    h.nnz = nnz;
    h.d = g.d;
    h.kind = kind;
    h.seed = g.seed;
    h.p_threshold = g.model == PPA_MODEL_BA ? (1ull << 31) : g.p_threshold;
    h.model = g.model;
    return h;
}

namespace {

// Degree histogram of a generated graph on the device: ppa_degree_histogram
// for the degrees and their maximum, then ppa_histogram_of_degrees into a
// histogram of that length.
// seedDeg (nullable) receives the degrees of nodes 0..d-1.
int deviceDegreeHistogram(ppa_ctx *ctx, const ppa_graph_desc &g, const int32_t *dGraph,
                          std::vector<uint64_t> *hist, std::vector<uint32_t> *seedDeg)
{
    uint32_t *dDeg = nullptr;
    uint64_t *dMax = nullptr, *dHist = nullptr;
    int rc = PPA_OK;
    if (hipMalloc(&dDeg, g.n * sizeof(uint32_t)) != hipSuccess || hipMalloc(&dMax, sizeof(uint64_t)) != hipSuccess)
        rc = PPA_ERR_NOMEM;
    uint64_t mx = 0;
    if (!rc) rc = ppa_degree_histogram(ctx, &g, dGraph, dDeg, nullptr, 0, dMax, nullptr);
    if (!rc && hipMemcpy(&mx, dMax, sizeof(mx), hipMemcpyDeviceToHost) != hipSuccess) rc = PPA_ERR_DEVICE;
    if (!rc) hist->assign(mx + 1, 0);
    if (!rc && hipMalloc(&dHist, hist->size() * sizeof(uint64_t)) != hipSuccess) rc = PPA_ERR_NOMEM;
    if (!rc) rc = ppa_histogram_of_degrees(ctx, g.n, dDeg, dHist, hist->size(), dMax, nullptr);
    if (!rc && hipMemcpy(hist->data(), dHist, hist->size() * sizeof(uint64_t), hipMemcpyDeviceToHost) != hipSuccess)
        rc = PPA_ERR_DEVICE;
    if (!rc && seedDeg) {
        seedDeg->assign(g.d, 0);
        if (hipMemcpy(seedDeg->data(), dDeg, g.d * sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess)
            rc = PPA_ERR_DEVICE;
    }
    if (dDeg) (void)hipFree(dDeg);
    if (dMax) (void)hipFree(dMax);
    if (dHist) (void)hipFree(dHist);
    return rc;
}

// --edges-file, --csr-file (unless csrDone) and --in-csr-file of a graph
// held on the device.
int writeDeviceGraphFiles(const ExperimentConfiguration &conf, ppa_ctx *ctx, const ppa_graph_desc &g,
                          const int32_t *dGraph, const int64_t *dRowPtr, bool csrDone = false)
{
    int rc = PPA_OK;
    if (!conf._edgesFile.empty()) {
        std::vector<int32_t> host(g.n * g.d);
        if (hipMemcpy(host.data(), dGraph, host.size() * sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess)
            return PPA_ERR_DEVICE;
        if (!writeEdgeList(conf._edgesFile, host.data(), g.n, g.d, g.model)) return PPA_ERR_INVALID;
    }
    if (!csrDone && !conf._csrFile.empty() &&
        !writeCsr(conf._csrFile, csrHeader(g, g.n * g.d, 0), dRowPtr, dGraph, true))
        return PPA_ERR_INVALID;
    if (!conf._inCsrFile.empty()) {
        int64_t *dPtr = nullptr;
        int32_t *dSrc = nullptr;
        rc = hipMalloc(&dPtr, (g.n + 1) * sizeof(int64_t)) == hipSuccess &&
                     hipMalloc(&dSrc, g.n * g.d * sizeof(int32_t)) == hipSuccess
                 ? PPA_OK
                 : PPA_ERR_NOMEM;
        if (!rc) rc = ppa_in_csr(ctx, &g, dGraph, dPtr, dSrc, nullptr);
        int64_t nnz = 0;
        if (!rc && hipMemcpy(&nnz, dPtr + g.n, sizeof(nnz), hipMemcpyDeviceToHost) != hipSuccess)
            rc = PPA_ERR_DEVICE;
        if (!rc && !writeCsr(conf._inCsrFile, csrHeader(g, (uint64_t)nnz, 1), dPtr, dSrc, true))
            rc = PPA_ERR_INVALID;
        if (dPtr) (void)hipFree(dPtr);
        if (dSrc) (void)hipFree(dSrc);
    }
    return rc;
}

}  // namespace

// ---- cuPPA ------------------------------------------------------------------

cuPPA::cuPPA(int64_t nNodes, uint32_t mEdgesPerNode, double p) : _N(nNodes), _M(mEdgesPerNode), _p(p) {}

cuPPA::~cuPPA()
{
    if (_dGraph) (void)hipFree(_dGraph);
    if (_dRowPtr) (void)hipFree(_dRowPtr);
    if (_ctx) ppa_ctx_destroy(_ctx);
}

ppa_graph_desc cuPPA::desc(const ExperimentConfiguration &conf) const
{
    ppa_graph_desc g{};
    g.n = (uint64_t)_N;
    g.d = _M;
    g.p_threshold = ppa_p_threshold(_p);
    g.seed = conf._seed;
    return g;
}

int cuPPA::generateGraph(const ExperimentConfiguration &conf)
{
    return run(conf, "StepwiseTiming-Original.csv", false);
}

// cuPPA_2.cu:181-198 generates one graph stage by stage and, after each
// stage whose queues stayed below half capacity, doubles both its thread
// count and the next stage's size (up to the configured maxima). Here the
// stages are consecutive node ranges of one graph (ppa_generate_range), each
// a persistent launch: the first runs on 1/8 of the full grid, and after each
// range whose time per generated node beat the previous range's by 3 % or
// more, the grid and the next range double (up to the full grid and n/8
// nodes). Every launch shape gives the same graph, so the tuning changes only
// the time. One StepwiseTiming row per range (cuPPA_2.cu:204-211);
// `Computation Time` is the whole graph's wall time.
int cuPPA::generateGraphAdaptive(const ExperimentConfiguration &conf)
{
    return run(conf, "StepwiseTiming-Adaptive.csv", true);
}

int cuPPA::generateAdaptiveRanges(const ppa_graph_desc &g, const ppa_tuning &tu, std::vector<ppa_stats> *rows,
                                  std::vector<uint64_t> *range_nodes, double *wall_ms)
{
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
        cus = 1;
    // the library's single-call grid: 16 wavefronts per CU
    const uint32_t threads = tu.block_threads ? tu.block_threads : 256;
    const uint64_t full = (uint64_t)cus * std::max<uint32_t>(1, 1024 / threads);
    uint64_t grid = std::max<uint64_t>(1, full / 8);
    const uint64_t cap = std::max<uint64_t>(1ull << 16, g.n / 8);
    uint64_t size = std::min<uint64_t>(cap, std::max<uint64_t>(1ull << 16, g.n / 64));
    bool final_tuning = false;
    double prev = -1.0;
    uint64_t u = 0;
    const auto t0 = std::chrono::steady_clock::now();
    int rc = PPA_OK;
    while (u < g.n) {
        uint64_t last = std::min<uint64_t>(g.n, u + size);
        if (g.n - last < size / 2) last = g.n;  // no short tail range
        ppa_tuning t = tu;
        t.max_blocks = (uint32_t)grid;
        if ((rc = ppa_ctx_set_tuning(_ctx, &t)) != PPA_OK) break;
        if ((rc = ppa_generate_range(_ctx, &g, _dGraph, _dRowPtr, u, last, nullptr)) != PPA_OK) break;
        if ((rc = ppa_ctx_sync(_ctx, nullptr)) != PPA_OK) break;
        ppa_stats s{};
        ppa_get_stats(_ctx, &s);
        rows->push_back(s);
        range_nodes->push_back(last - u);
        const uint64_t lo = std::max<uint64_t>(u, (uint64_t)g.d + 1);
        const double per_node = s.kernel_ms / (double)std::max<uint64_t>(1, last > lo ? last - lo : 0);
        if (!final_tuning && (prev < 0.0 || per_node < 0.97 * prev)) {
            grid = std::min<uint64_t>(full, grid * 2);
            size = std::min<uint64_t>(cap, size * 2);
            if (grid == full && size == cap) final_tuning = true;
        }
        prev = per_node;
        u = last;
    }
    *wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    // later calls on this context keep the grid the tuning reached
    ppa_tuning kept = tu;
    kept.max_blocks = grid >= full ? 0u : (uint32_t)grid;
    const int r2 = ppa_ctx_set_tuning(_ctx, &kept);
    return rc ? rc : r2;
}

int cuPPA::run(const ExperimentConfiguration &conf, const char *timingCsv, bool adaptive)
{
    const ppa_graph_desc g = desc(conf);
    int rc = ppa_check_desc(&g);
    if (rc) return rc;
    if (!_ctx && (rc = ppa_ctx_create(&_ctx, conf._device, 0)) != PPA_OK) return rc;
    ppa_tuning tu{};
    if (conf._threadsGiven) {
        const int threads = conf._threadsPerBlock < 64 ? 64 : conf._threadsPerBlock;
        tu.block_threads = (uint32_t)((threads + 63) / 64 * 64);
        if (tu.block_threads > 256) tu.block_threads = 256;
    }
    if (conf._blocksGiven && conf._blockSize > 0) tu.max_blocks = (uint32_t)conf._blockSize;
    tu.chunk_nodes = conf._chunk;
    if ((rc = ppa_ctx_set_tuning(_ctx, &tu)) != PPA_OK) return rc;

    if (!_dGraph) {
        if (hipMalloc(&_dGraph, g.n * g.d * sizeof(int32_t)) != hipSuccess) return PPA_ERR_NOMEM;
        if (hipMalloc(&_dRowPtr, (g.n + 1) * sizeof(int64_t)) != hipSuccess) return PPA_ERR_NOMEM;
    }
    std::printf("--------------------------------------------------------------------------------\n");
    std::printf("Graph: %.2f MB\t| RowPtr: %.2f MB\n", g.n * g.d * 4.0 / 1048576.0, (g.n + 1) * 8.0 / 1048576.0);
    // with --csr-file (and one launch shape) the file is written while the graph is generated
    const bool ranged = adaptive && !conf._blocksGiven;
    const bool streamCsr = !conf._csrFile.empty() && !ranged;
    ppa_stats st{};
    std::vector<ppa_stats> stages;        // one StepwiseTiming row each
    std::vector<uint64_t> stage_nodes;
    double computation_ms = 0.0;
    if (ranged) {
        rc = generateAdaptiveRanges(g, tu, &stages, &stage_nodes, &computation_ms);
        for (const ppa_stats &s : stages) {
            st.draws += s.draws;
            st.retries += s.retries;
            st.prefetch_misses += s.prefetch_misses;
            st.grid_blocks = s.grid_blocks;  // the grid the tuning ended at
            st.block_threads = s.block_threads;
        }
    } else {
        if (streamCsr) {
            rc = generateStreamingCsr(conf, g, &st);
        } else {
            rc = ppa_generate(_ctx, &g, _dGraph, _dRowPtr, nullptr);
            ppa_get_stats(_ctx, &st);
        }
        stages.push_back(st);
        stage_nodes.push_back(g.n);
        computation_ms = st.total_ms;
    }
    if (rc) return rc;
    std::printf("BLOCKS: %llu\t| THREADS PER BLOCK: %llu\t| DRAWS: %llu\t| RETRIES: %llu\t| PARKED: %llu\n",
                (unsigned long long)st.grid_blocks, (unsigned long long)st.block_threads,
                (unsigned long long)st.draws, (unsigned long long)st.retries,
                (unsigned long long)st.prefetch_misses);
    {
        // cuPPA_1.cu:84-87,217-224 and cuPPA_2.cu:204-211: one row per stage
        // (the adaptive mode's node ranges, the plain mode's single launch).
        std::ofstream csv(timingCsv);
        csv << "Round,nBlocks,nThreads,Edges,MaxQ,Initialization Time,Computation Time\n";
        const size_t rows = std::min<size_t>(stages.size(), conf._N_STAGE_TIMINGS > 0 ? conf._N_STAGE_TIMINGS : 0);
        for (size_t i = 0; i < rows; ++i) {
            const ppa_stats &x = stages[i];
            csv << i + 1 << "," << x.grid_blocks << "," << x.block_threads << "," << stage_nodes[i] * g.d << ","
                << x.prefetch_misses << "," << (x.total_ms - x.kernel_ms) * 1e-3 << "," << x.kernel_ms * 1e-3
                << "\n";
        }
    }
    std::printf("Computation Time: %f\n\n\n", computation_ms * 1e-3);
    if (conf._FLAG_SHOW_DEGREE_DISTRIBUTION && (rc = printDegreeDistribution(conf)) != PPA_OK) return rc;
    if ((rc = writeDeviceGraphFiles(conf, _ctx, g, _dGraph, _dRowPtr, streamCsr)) != PPA_OK) return rc;
    std::cout << "Ended" << std::endl;
    return PPA_OK;
}

int cuPPA::generateStreamingCsr(const ExperimentConfiguration &conf, const ppa_graph_desc &g, ppa_stats *total)
{
    FILE *f = std::fopen(conf._csrFile.c_str(), "wb");
    if (!f) return PPA_ERR_INVALID;
    const CsrHeader h = csrHeader(g, g.n * g.d, 0);
    bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1;
    DeviceWriter w;
    hipStream_t gen = nullptr;
    int rc = hipStreamCreateWithFlags(&gen, hipStreamNonBlocking) == hipSuccess ? PPA_OK : PPA_ERR_DEVICE;
    // eight ranges, of at least 2^20 nodes
    const uint64_t step = std::max<uint64_t>(1ull << 20, (g.n + 7) / 8);
    std::vector<uint64_t> bounds{0};
    while (bounds.back() < g.n) bounds.push_back(std::min<uint64_t>(g.n, bounds.back() + step));
    *total = ppa_stats{};
    auto finish = [&]() {
        int r = ppa_ctx_sync(_ctx, gen);
        ppa_stats s{};
        ppa_get_stats(_ctx, &s);
        total->total_ms += s.total_ms;
        total->kernel_ms += s.kernel_ms;
        total->draws += s.draws;
        total->retries += s.retries;
        total->prefetch_misses += s.prefetch_misses;
        total->grid_blocks = std::max(total->grid_blocks, s.grid_blocks);
        total->block_threads = s.block_threads;
        return r;
    };
    if (!rc) rc = ppa_generate_range(_ctx, &g, _dGraph, _dRowPtr, 0, bounds[1], gen);
    if (!rc) rc = finish();
    for (size_t k = 1; !rc && ok && k < bounds.size(); ++k) {
        const bool more = k + 1 < bounds.size();
        if (more) rc = ppa_generate_range(_ctx, &g, _dGraph, _dRowPtr, bounds[k], bounds[k + 1], gen);
        // while range k runs: the row offsets (filled by range 0), then range k-1's rows
        if (!rc && k == 1) ok = w.write(f, _dRowPtr, (g.n + 1) * sizeof(int64_t));
        if (!rc && ok)
            ok = w.write(f, _dGraph + bounds[k - 1] * g.d, (bounds[k] - bounds[k - 1]) * g.d * sizeof(int32_t));
        if (more) {
            const int r = finish();  // always ends a launched range, whatever the writes did
            if (!rc) rc = r;
        }
    }
    if (gen) (void)hipStreamDestroy(gen);
    ok = std::fclose(f) == 0 && ok;
    return rc ? rc : (ok ? PPA_OK : PPA_ERR_INVALID);
}

int cuPPA::printDegreeDistribution(const ExperimentConfiguration &conf)
{
    if (!_dGraph || !_ctx) return PPA_ERR_INVALID;
    std::printf("Generating Degree Distribution\n");
    const ppa_graph_desc g = desc(conf);
    std::vector<uint64_t> hist;
    std::vector<uint32_t> seedDeg;
    int rc = deviceDegreeHistogram(_ctx, g, _dGraph, &hist, &seedDeg);
    if (rc) return rc;
    // the reference prints the degrees of its seed nodes (cuPPA.cu:136-139)
    for (uint32_t u = 0; u < _M; ++u) std::cout << u << "\t" << seedDeg[u] << std::endl;
    return writeDegreeCsv(joinPath(conf._outputDir, conf._distfile1), hist) ? PPA_OK : PPA_ERR_INVALID;
}

// ---- SerialBA -----------------------------------------------------------------

SerialBA::~SerialBA()
{
    if (_dGraph) (void)hipFree(_dGraph);
    if (_dRowPtr) (void)hipFree(_dRowPtr);
    if (_ctx) ppa_ctx_destroy(_ctx);
}

int SerialBA::generateGraph(const ExperimentConfiguration &conf)
{
    const std::string METHOD = "SBA";
    std::cout << METHOD << std::endl;
    std::ofstream log(conf._outputDir + METHOD + ".log");  // SerialBA.cpp:72-81
    log << METHOD << std::endl << _N << std::endl << _M << std::endl;
    ppa_graph_desc g{};
    g.n = (uint64_t)_N;
    g.d = _M;
    g.model = PPA_MODEL_BA;
    g.seed = conf._seed;
    int rc = ppa_check_desc(&g);
    if (rc) return rc;
    // SerialBA.cpp:24-37: the graph's size, then the allocation
    log << "Required Memory for Graph: " << (g.n * g.d * 4.0 + (g.n + 1) * 8.0) / (1024.0 * 1024.0) << std::endl;
    if (!_ctx && (rc = ppa_ctx_create(&_ctx, conf._device, 0)) != PPA_OK) return rc;
    if (!_dGraph) {
        if (hipMalloc(&_dGraph, g.n * g.d * sizeof(int32_t)) != hipSuccess) {
            log << "Not Enough Memory!" << std::endl;
            return PPA_ERR_NOMEM;
        }
        if (hipMalloc(&_dRowPtr, (g.n + 1) * sizeof(int64_t)) != hipSuccess) return PPA_ERR_NOMEM;
    }
    log << "Memory Allocated" << std::endl;
    if ((rc = ppa_generate(_ctx, &g, _dGraph, _dRowPtr, nullptr)) != PPA_OK) return rc;
    ppa_stats st{};
    ppa_get_stats(_ctx, &st);
    // SerialBA.cpp:157-162
    std::cout << "Computation Time : " << st.total_ms * 1e-3 << std::endl;
    log << "Computation Time : " << st.total_ms * 1e-3 << std::endl;
    if (conf._FLAG_SHOW_DEGREE_DISTRIBUTION) {
        // PAHelper.hpp:79-128
        std::cout << "MAXIMUM SIZE: " << g.n * g.d * 2 << std::endl;
        std::vector<uint64_t> hist;
        if ((rc = deviceDegreeHistogram(_ctx, g, _dGraph, &hist, nullptr)) != PPA_OK) return rc;
        std::printf("Degree of each vertex is computed!\n");
        if (!writeDegreeCsv(joinPath(conf._outputDir, conf._distfile1), hist)) return PPA_ERR_INVALID;
    }
    return writeDeviceGraphFiles(conf, _ctx, g, _dGraph, _dRowPtr);
}

// ---- SerialCopyModel ----------------------------------------------------------

int SerialCopyModel::generateGraph(const ExperimentConfiguration &conf)
{
    const std::string METHOD = "SCM";
    std::cout << METHOD << std::endl;
    std::ofstream log(conf._outputDir + METHOD + ".log");  // SerialCopyModel3.cpp:19-25
    log << METHOD << std::endl << _N << std::endl << _M << std::endl;
    ppa_graph_desc g{};
    g.n = (uint64_t)_N;
    g.d = _M;
    g.p_threshold = ppa_p_threshold(_p);
    g.seed = conf._seed;
    int rc = ppa_check_desc(&g);
    if (rc) return rc;
    _G.assign(g.n * g.d, -1);
    const auto t0 = std::chrono::steady_clock::now();
    rc = ppa_cpu_generate(&g, _G.data(), nullptr, nullptr);
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rc) return rc;
    std::cout << "Computation Time : " << secs << std::endl;
    log << "Computation Time : " << secs << std::endl;
    if (conf._FLAG_SHOW_DEGREE_DISTRIBUTION) {
        std::vector<uint32_t> deg(g.n);
        uint64_t mx = 0;
        if ((rc = ppa_cpu_degree_histogram(&g, _G.data(), deg.data(), nullptr, 0, &mx)) != PPA_OK) return rc;
        std::vector<uint64_t> hist(mx + 1, 0);
        if ((rc = ppa_cpu_degree_histogram(&g, _G.data(), deg.data(), hist.data(), hist.size(), &mx)) != PPA_OK)
            return rc;
        if (!writeDegreeCsv(joinPath(conf._outputDir, conf._distfile1), hist)) return PPA_ERR_INVALID;
    }
    if (!conf._edgesFile.empty() && !writeEdgeList(conf._edgesFile, _G.data(), g.n, g.d)) return PPA_ERR_INVALID;
    if (!conf._csrFile.empty()) {
        std::vector<int64_t> rowPtr(g.n + 1);
        for (uint64_t t = 0; t <= g.n; ++t) rowPtr[t] = (int64_t)(t * g.d);
        if (!writeCsr(conf._csrFile, csrHeader(g, g.n * g.d, 0), rowPtr.data(), _G.data(), false))
            return PPA_ERR_INVALID;
    }
    return PPA_OK;
}
