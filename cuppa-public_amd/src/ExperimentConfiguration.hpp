// ExperimentConfiguration.hpp -- parameters of a generation run.
//
// Keeps the field names and defaults of the reference's ExperimentConfiguration
// (src/Utility/ExperimentConfiguration.hpp:51-97) and of parseExperiments
// (src/main.cu:129-178), the command-line flags of src/main.cu:182-218 and the
// [config] INI keys of ExperimentConfiguration.hpp:114-189. The parser and the
// INI reader are this project's own (the reference vendors CLI11 and minIni).
// Added: --seed (Philox key; the reference seeds with time(0), cuPPA_1.cu:39),
// --device, --chunk, --edges-file, --csr-file and --in-csr-file.
#pragma once

#include <cstdint>
#include <string>

struct ExperimentConfiguration {
    std::string _algorithm;

    int64_t _n1 = 1000;
    uint32_t _x1 = 4;
    uint32_t _d1min = 0, _d1max = 0, _d1avg = 0;
    std::string _weight1 = "none";
    std::string _distfile1 = "degreeDistribution.txt";
    std::string _avgcc1, _probabilityMatrixFile1;

    int64_t _n2 = 1000;
    uint32_t _x2 = 4;
    uint32_t _d2min = 0, _d2max = 0, _d2avg = 0;
    std::string _weight2 = "none";
    std::string _distfile2 = "degreeDistribution.txt";
    std::string _avgcc2, _probabilityMatrixFile2;

    long _msgBufferSize = 1024;
    long _mpiBufferSize = 1024;

    int _blockSize = 2;        // -b: reference workgroups; here a cap on the persistent grid
    int _threadsPerBlock = 32; // -t: threads per workgroup (rounded up to a wavefront multiple)
    int _warpSize = 32;        // -w: accepted; gfx950 wavefronts are 64 wide

    long _stages = 1000;       // -s: accepted; one persistent launch covers every stage
    double _lbFactor = 1;
    double _alpha = 0.5;       // p, probability of a direct edge
    double _gamma = 2.7;
    double _probability = 0.1;
    std::string _loadBalancing = "none";
    std::string _outputDir = "./";
    int _write = 0;

    bool _FLAG_SHOW_DEGREE_DISTRIBUTION = false;
    int _N_STAGE_TIMINGS = 0;

    // additions of this implementation
    bool _blocksGiven = false;   // -b/-t were on the command line: use them as
    bool _threadsGiven = false;  // a grid cap / workgroup size, else fill the GPU
    uint64_t _seed = 0;
    int _device = 0;
    uint32_t _chunk = 0;
    std::string _edgesFile;
    std::string _csrFile;    // binary out-edge CSR (row offsets + targets)
    std::string _inCsrFile;  // binary in-edge CSR (GPU paths)

    // INI loader: every key of the [config] section with the reference's
    // defaults for missing keys (ExperimentConfiguration.hpp:114-189).
    bool load(const std::string &inifile, std::string *err);
};

// Parses argv like src/main.cu:123-236 (INI, if given, is applied after the
// flags and overrides them, as in the reference). Returns 0 on success, 1 on
// a parse error (message in *err), 2 when --help was printed.
int parseExperiments(int argc, char **argv, ExperimentConfiguration *out, std::string *err);
void printUsage(const char *prog);
