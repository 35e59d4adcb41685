// main.cpp -- command-line driver (counterpart of src/main.cu:46-260).
//
//   ppa_runner -A cuppa|cuppa-adaptive|scm|sba -n N -x M -a p [--seed S] [--show-dd 1] ...
//
// Dispatch follows syntheticNetworkGeneration (main.cu:91-119); an unknown
// algorithm is reported as "Invalid Algorithm" (ExperimentConfiguration.hpp:
// 390-399) and ends the run with exit status 1.
#include <cstdio>
#include <iostream>
#include <string>

#include "ExperimentConfiguration.hpp"
#include "Generators.hpp"

namespace {

void printConfig(const ExperimentConfiguration &c)
{
    // same fields as main.cu:50-69
    std::printf("Method:\t%s\tLoad Balancing:\t%s\tWeight:\t%s\n", c._algorithm.c_str(), c._loadBalancing.c_str(),
                c._weight1.c_str());
    std::printf("N:\t%lld\tM:\t%u\n", (long long)c._n1, c._x1);
    std::printf("Min Degree:\t%u\tMax Degree:\t%u\tAvg Degree:\t%u\n", c._d1min, c._d1max, c._d1avg);
    std::printf("Output:\t%s\n", c._outputDir.c_str());
    std::printf("Probability:\t%f\tProcessors:\t%d\tGamma:\t%f\ta:\t%f\n", c._probability, 1, c._gamma, c._alpha);
    std::printf("MSG Buffer:\t%ld\tMPI Buffer:\t%ld\tLB Factor:\t%f\tNum Stages:\t%ld\tWrite To File:\t%d\n",
                c._msgBufferSize, c._mpiBufferSize, c._lbFactor, c._stages, c._write);
    std::printf("Seed:\t%llu\n", (unsigned long long)c._seed);
}

int report(int rc)
{
    if (rc != PPA_OK) std::fprintf(stderr, "error: %s (%d)\n", ppa_strerror(rc), rc);
    return rc == PPA_OK ? 0 : 1;
}

}  // namespace

int main(int argc, char **argv)
{
    std::printf("cuPPA copy-model generator (MI355X / gfx950)\n");
    ExperimentConfiguration conf;
    std::string err;
    const int prc = parseExperiments(argc, argv, &conf, &err);
    if (prc == 2) return 0;
    if (prc != 0) {
        std::cout << "Exception" << err << std::endl;  // main.cu:223-227
        return 1;
    }
    printConfig(conf);
    if (conf._algorithm == "scm") {
        SerialCopyModel scm(conf._n1, conf._x1, conf._alpha);
        return report(scm.generateGraph(conf));
    }
    if (conf._algorithm == "sba") {
        SerialBA ba(conf._n1, conf._x1);
        return report(ba.generateGraph(conf));
    }
    if (conf._algorithm == "cuppa" || conf._algorithm == "cuppa-adaptive") {
        cuPPA cuppa(conf._n1, conf._x1, conf._alpha);
        return report(conf._algorithm == "cuppa" ? cuppa.generateGraph(conf) : cuppa.generateGraphAdaptive(conf));
    }
    std::cout << "Algorithm: " << conf._algorithm << std::endl;
    std::cerr << "Invalid Algorithm" << std::endl;
    return 1;
}
