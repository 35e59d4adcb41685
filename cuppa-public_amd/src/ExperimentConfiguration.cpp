// ExperimentConfiguration.cpp -- flag and INI parsing for the host driver.
#include "ExperimentConfiguration.hpp"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <functional>
#include <map>
#include <sstream>

namespace {

std::string trim(const std::string &s)
{
    const char *ws = " \t\r\n";
    const size_t a = s.find_first_not_of(ws);
    if (a == std::string::npos) return "";
    const size_t b = s.find_last_not_of(ws);
    return s.substr(a, b - a + 1);
}

// Reads the [section] keys of an INI file into a map (later keys win).
bool read_ini(const std::string &path, const std::string &section, std::map<std::string, std::string> *kv,
              std::string *err)
{
    std::ifstream in(path);
    if (!in) {
        if (err) *err = "cannot open config file " + path;
        return false;
    }
    std::string line, cur;
    while (std::getline(in, line)) {
        line = trim(line);
        if (line.empty() || line[0] == ';' || line[0] == '#') continue;
        if (line.front() == '[' && line.back() == ']') {
            cur = trim(line.substr(1, line.size() - 2));
            continue;
        }
        if (cur != section) continue;
        const size_t eq = line.find_first_of("=:");
        if (eq == std::string::npos) continue;
        std::string key = trim(line.substr(0, eq)), val = trim(line.substr(eq + 1));
        if (val.size() >= 2 && val.front() == '"' && val.back() == '"') val = val.substr(1, val.size() - 2);
        (*kv)[key] = val;
    }
    return true;
}

bool parse_bool(const std::string &v)
{
    return v == "1" || v == "true" || v == "TRUE" || v == "True" || v == "yes" || v == "on";
}

}  // namespace

bool ExperimentConfiguration::load(const std::string &inifile, std::string *err)
{
    std::map<std::string, std::string> kv;
    if (!read_ini(inifile, "config", &kv, err)) return false;
    auto get = [&](const char *key, const char *def) {
        auto it = kv.find(key);
        return it == kv.end() ? std::string(def) : it->second;
    };
    _algorithm = get("algorithm", "");
    _n1 = std::atoll(get("n1", "1000").c_str());
    _x1 = (uint32_t)std::atoi(get("x1", "2").c_str());
    _d1min = (uint32_t)std::atoi(get("d1min", "2").c_str());
    _d1max = (uint32_t)std::atoi(get("d1max", "10").c_str());
    _d1avg = (uint32_t)std::atoi(get("d1avg", "5").c_str());
    _weight1 = get("weight1", "none");
    _distfile1 = get("distfile1", "");
    _avgcc1 = get("avgccfile1", "");
    _probabilityMatrixFile1 = get("pmatfile1", "");
    _alpha = std::atof(get("alpha", "0.5").c_str());
    _gamma = std::atof(get("gamma", "2.7").c_str());
    _n2 = std::atoll(get("n2", "0").c_str());
    _x2 = (uint32_t)std::atoi(get("x2", "2").c_str());
    _d2min = (uint32_t)std::atoi(get("d2min", "2").c_str());
    _d2max = (uint32_t)std::atoi(get("d2max", "10").c_str());
    _d2avg = (uint32_t)std::atoi(get("d2avg", "5").c_str());
    _weight2 = get("weight2", "none");
    _distfile2 = get("distfile2", "");
    _avgcc2 = get("avgccfile2", "");
    _probabilityMatrixFile2 = get("pmatfile2", "");
    _outputDir = get("outputdir", "./out/");
    _msgBufferSize = std::atoi(get("msg_buffer", "1024").c_str());
    _mpiBufferSize = std::atoi(get("mpi_buffer", "1024").c_str());
    _lbFactor = std::atof(get("lbfactor", "1.0").c_str());
    _probability = std::atof(get("probability", "0.1").c_str());
    _stages = (long)std::atof(get("stages", "1000").c_str());
    _loadBalancing = get("loadbalancing", "none");
    _write = std::atoi(get("write", "0").c_str());
    // keys of this implementation (absent from the reference's INI schema)
    _seed = std::strtoull(get("seed", std::to_string(_seed).c_str()).c_str(), nullptr, 0);
    return true;
}

void printUsage(const char *prog)
{
    std::printf(
        "usage: %s [options]\n"
        "  -c,--config-file FILE     INI file, section [config] (applied after the flags)\n"
        "  -A,--algorithm NAME       cuppa | cuppa-adaptive | scm | sba\n"
        "  -l,--loadbalancing NAME   accepted (rrp is the only scheme of the reference)\n"
        "  -n,--n1 N                 number of vertices\n"
        "  -x,--x1 M                 out-edges per new vertex (1..8192)\n"
        "  --weight1 W, -N,--n2 N, -X,--x2 M, --weight2 W   accepted, unused\n"
        "  -a,--alpha P              probability of creating a direct edge\n"
        "  -s,--stages S             accepted (one persistent launch covers all stages)\n"
        "  --msg-buffer-size B, --mpi-buffer-size B          accepted, unused\n"
        "  -b,--blocks B             cap on persistent workgroups (0 = fill the GPU)\n"
        "  -t,--threads T            threads per workgroup (rounded up to 64, <= 256)\n"
        "  -w,--warps W              accepted (wavefronts are 64 wide)\n"
        "  --show-dd 0|1             write the Degree,Count CSV to <outputdir>/<dd-filename>\n"
        "  --dd-filename NAME        degree distribution file name\n"
        "  --show-time-n-stages K    rows of StepwiseTiming CSV to write\n"
        "  --seed S                  Philox key (default 0)\n"
        "  --device D                GPU ordinal\n"
        "  --chunk C                 nodes per wavefront claim (multiple of 64)\n"
        "  --edges-file PATH         write the edge list (\"u v\" lines)\n"
        "  --csr-file PATH           write the out-edge CSR (binary, see writeCsr)\n"
        "  --in-csr-file PATH        write the in-edge CSR (binary; cuppa paths)\n"
        "  -o,--output-dir DIR       output directory\n",
        prog);
}

int parseExperiments(int argc, char **argv, ExperimentConfiguration *e, std::string *err)
{
    std::string conf_ini;
    using Setter = std::function<bool(const std::string &)>;
    auto i64 = [](int64_t *dst) { return Setter([dst](const std::string &v) { char *end; *dst = std::strtoll(v.c_str(), &end, 10); return *end == 0; }); };
    auto u32 = [](uint32_t *dst) { return Setter([dst](const std::string &v) { char *end; long long x = std::strtoll(v.c_str(), &end, 10); *dst = (uint32_t)x; return *end == 0 && x >= 0 && x <= 0xffffffffLL; }); };
    auto u64 = [](uint64_t *dst) { return Setter([dst](const std::string &v) { char *end; *dst = std::strtoull(v.c_str(), &end, 0); return *end == 0; }); };
    auto lng = [](long *dst) { return Setter([dst](const std::string &v) { char *end; *dst = std::strtol(v.c_str(), &end, 10); return *end == 0; }); };
    auto i32 = [](int *dst) { return Setter([dst](const std::string &v) { char *end; *dst = (int)std::strtol(v.c_str(), &end, 10); return *end == 0; }); };
    auto dbl = [](double *dst) { return Setter([dst](const std::string &v) { char *end; *dst = std::strtod(v.c_str(), &end); return *end == 0; }); };
    auto str = [](std::string *dst) { return Setter([dst](const std::string &v) { *dst = v; return true; }); };
    auto bln = [](bool *dst) { return Setter([dst](const std::string &v) { *dst = parse_bool(v); return true; }); };
    auto seen = [](Setter s, bool *flag) { return Setter([s, flag](const std::string &v) { *flag = true; return s(v); }); };

    const std::map<std::string, Setter> opts = {
        {"-c", str(&conf_ini)}, {"--config-file", str(&conf_ini)},
        {"-A", str(&e->_algorithm)}, {"--algorithm", str(&e->_algorithm)},
        {"-l", str(&e->_loadBalancing)}, {"--loadbalancing", str(&e->_loadBalancing)},
        {"-n", i64(&e->_n1)}, {"--n1", i64(&e->_n1)},
        {"-x", u32(&e->_x1)}, {"--x1", u32(&e->_x1)},
        {"--weight1", str(&e->_weight1)},
        {"-N", i64(&e->_n2)}, {"--n2", i64(&e->_n2)},
        {"-X", u32(&e->_x2)}, {"--x2", u32(&e->_x2)},
        {"--weight2", str(&e->_weight2)},
        {"-a", dbl(&e->_alpha)}, {"--alpha", dbl(&e->_alpha)},
        {"-s", lng(&e->_stages)}, {"--stages", lng(&e->_stages)},
        {"--msg-buffer-size", lng(&e->_msgBufferSize)}, {"--mpi-buffer-size", lng(&e->_mpiBufferSize)},
        {"-b", seen(i32(&e->_blockSize), &e->_blocksGiven)},
        {"--blocks", seen(i32(&e->_blockSize), &e->_blocksGiven)},
        {"-t", seen(i32(&e->_threadsPerBlock), &e->_threadsGiven)},
        {"--threads", seen(i32(&e->_threadsPerBlock), &e->_threadsGiven)},
        {"-w", i32(&e->_warpSize)}, {"--warps", i32(&e->_warpSize)},
        {"--show-dd", bln(&e->_FLAG_SHOW_DEGREE_DISTRIBUTION)},
        {"--dd-filename", str(&e->_distfile1)},
        {"--show-time-n-stages", i32(&e->_N_STAGE_TIMINGS)},
        {"--seed", u64(&e->_seed)},
        {"--device", i32(&e->_device)},
        {"--chunk", u32(&e->_chunk)},
        {"--edges-file", str(&e->_edgesFile)},
        {"--csr-file", str(&e->_csrFile)},
        {"--in-csr-file", str(&e->_inCsrFile)},
        {"-o", str(&e->_outputDir)}, {"--output-dir", str(&e->_outputDir)},
    };
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "-h" || a == "--help") {
            printUsage(argv[0]);
            return 2;
        }
        std::string val;
        bool have_val = false;
        const size_t eq = a.find('=');
        if (a.rfind("--", 0) == 0 && eq != std::string::npos) {
            val = a.substr(eq + 1);
            a = a.substr(0, eq);
            have_val = true;
        }
        auto it = opts.find(a);
        if (it == opts.end()) {
            if (err) *err = "The following argument was not expected: " + a;
            return 1;
        }
        if (!have_val) {
            if (i + 1 >= argc) {
                if (err) *err = a + " requires an argument";
                return 1;
            }
            val = argv[++i];
        }
        if (!it->second(val)) {
            if (err) *err = "Could not convert: " + a + " = " + val;
            return 1;
        }
    }
    if (!conf_ini.empty() && !e->load(conf_ini, err)) return 1;
    return 0;
}
