// Generators.hpp -- host-side generator objects of the driver.
//
// cuPPA mirrors the reference class of src/PA/cuPPA/cuPPA.cuh:84-105 (same
// constructor arguments and entry points) over the C ABI of include/ppa.h;
// SerialCopyModel mirrors src/PA/SPA/SerialCopyModel.h:26-47 over
// ppa_cpu_generate; SerialBA mirrors src/PA/SPA/SerialBA.h:30-51 and runs the
// Barabasi-Albert model on the GPU (PPA_MODEL_BA). Unlike the reference, a
// generator keeps its graph until it is destroyed and reports failures as a
// nonzero return value.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "ExperimentConfiguration.hpp"
#include "ppa.h"

class cuPPA {
public:
    cuPPA(int64_t nNodes, uint32_t mEdgesPerNode, double p);
    ~cuPPA();
    cuPPA(const cuPPA &) = delete;
    cuPPA &operator=(const cuPPA &) = delete;

    // cuPPA_1.cu:9-254. Returns 0 or a PPA_ERR_* code.
    int generateGraph(const ExperimentConfiguration &conf);
    // cuPPA_2.cu:9-241: the reference grows threads and stage size online
    // while its queues stay half empty. Here one graph is generated as
    // consecutive node ranges (ppa_generate_range) starting on 1/8 of the
    // persistent grid; the grid and the next range double while the time per
    // node improves. One StepwiseTiming row per range; Computation Time is
    // the whole graph's. With -b given, one launch of that grid instead.
    int generateGraphAdaptive(const ExperimentConfiguration &conf);
    // cuPPA.cu:103-165
    int printDegreeDistribution(const ExperimentConfiguration &conf);

    const int32_t *deviceTargets() const { return _dGraph; }

private:
    int run(const ExperimentConfiguration &conf, const char *timingCsv, bool adaptive);
    // Generation in node ranges (ppa_generate_range) with --csr-file written
    // as they finish: range k's rows are copied out and written while range
    // k+1 is generated. *total sums the ranges' statistics.
    int generateStreamingCsr(const ExperimentConfiguration &conf, const ppa_graph_desc &g, ppa_stats *total);
    // The adaptive mode's ranges: their statistics, their node counts and the
    // wall time of the whole graph.
    int generateAdaptiveRanges(const ppa_graph_desc &g, const ppa_tuning &tu, std::vector<ppa_stats> *rows,
                               std::vector<uint64_t> *range_nodes, double *wall_ms);
    ppa_graph_desc desc(const ExperimentConfiguration &conf) const;

    int64_t _N;
    uint32_t _M;
    double _p;
    int32_t *_dGraph = nullptr;
    int64_t *_dRowPtr = nullptr;
    ppa_ctx *_ctx = nullptr;
};

class SerialCopyModel {
public:
    SerialCopyModel(int64_t n, uint32_t m, double p) : _N(n), _M(m), _p(p) {}
    // SerialCopyModel3.cpp:11-107
    int generateGraph(const ExperimentConfiguration &conf);
    const std::vector<int32_t> &targets() const { return _G; }

private:
    int64_t _N;
    uint32_t _M;
    double _p;
    std::vector<int32_t> _G;
};

class SerialBA {
public:
    SerialBA(int64_t n, uint32_t m = 1) : _N(n), _M(m) {}
    ~SerialBA();
    SerialBA(const SerialBA &) = delete;
    SerialBA &operator=(const SerialBA &) = delete;
    // SerialBA.cpp:68-187 (generateGraphNetworkX) with the degree CSV of
    // PAHelper::outputSBADegreeDistribution (PAHelper.hpp:69-147).
    int generateGraph(const ExperimentConfiguration &conf);

    const int32_t *deviceTargets() const { return _dGraph; }

private:
    int64_t _N;
    uint32_t _M;
    int32_t *_dGraph = nullptr;
    int64_t *_dRowPtr = nullptr;
    ppa_ctx *_ctx = nullptr;
};

// Output formats of the reference.
// "Degree,Count" CSV of cuPPA.cu:148-160 and PAHelper.hpp:54-63.
bool writeDegreeCsv(const std::string &path, const std::vector<uint64_t> &hist);
// "u v" edge list of PAHelper::writeGraphToDisk (PAHelper.hpp:163-177): the
// seed clique's edges once each, then every slot of nodes t >= d+1. For a
// Barabasi-Albert graph: every slot of nodes t >= d (the pairs of nodes u >= M
// that PAHelper.hpp:84-121 counts; the seed self-loops are left out).
bool writeEdgeList(const std::string &path, const int32_t *targets, uint64_t n, uint32_t d,
                   uint32_t model = PPA_MODEL_COPY);
std::string joinPath(const std::string &dir, const std::string &file);

// Binary CSR file (this project's format; the reference writes text only):
// a 64-byte little-endian header
//   char magic[8] = "PPACSR1"; uint64 n; uint64 nnz; uint32 d; uint32 kind
//   (0 = out-edges: row t lists t's targets; 1 = in-edges: row v lists the
//   sources naming v, ascending); uint64 seed; uint64 p_threshold; uint64 model
//   (PPA_MODEL_*); uint64 0
// then int64 offsets[n + 1] and int32 indices[nnz]. Arrays given as device
// pointers are streamed through pinned staging buffers, the copy of one chunk
// overlapping the write of the previous one.
struct CsrHeader {
    char magic[8];
    uint64_t n, nnz;
    uint32_t d, kind;
    uint64_t seed, p_threshold;
    uint64_t model;
    uint64_t reserved;
};
static_assert(sizeof(CsrHeader) == 64, "CSR header layout");
bool writeCsr(const std::string &path, const CsrHeader &h, const int64_t *offsets, const int32_t *indices,
              bool device);
