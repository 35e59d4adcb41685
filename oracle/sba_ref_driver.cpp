// Driver for the reference's own serial Barabasi-Albert generator (test
// infrastructure only).
//
// Compiled by `make -C oracle ref` together with the reference sources where
// they lie (src/PA/SPA/SerialBA.cpp, src/Common/Timer.cpp). It mirrors how
// src/main.cu:99-103 calls it: construct SerialBA(n, m) and call
// generateGraph(conf) with the degree-distribution flag set, so the reference
// writes its own "Degree,Count" CSV (PAHelper.hpp:69-147). Its draws come from
// MT19937-64 with the default seed (Random64.hpp:75-76), so its output is
// compared with ours statistically, not bit for bit.
//
// usage: sba_ref <n> <m> <outdir/> [write_dd=1]
#include <cstdlib>
#include <iostream>

#include "PA/SPA/SerialBA.h"

int main(int argc, char **argv)
{
    if (argc != 4 && argc != 5) {
        std::cerr << "usage: sba_ref <n> <m> <outdir/> [write_dd=1]" << std::endl;
        return 2;
    }
    ExperimentConfiguration conf;
    conf._n1 = (NodeType)std::atoll(argv[1]);
    conf._x1 = (WeightType)std::atoi(argv[2]);
    conf._outputDir = argv[3];
    conf._distfile1 = "sba_dd.csv";
    conf._FLAG_SHOW_DEGREE_DISTRIBUTION = argc == 5 ? std::atoi(argv[4]) != 0 : true;
    conf._N_STAGE_TIMINGS = 0;
    SerialBA ba(conf._n1, conf._x1);
    ba.generateGraph(conf);
    return 0;
}
