// Driver for the reference's own serial copy model (test infrastructure only).
//
// Compiled by `make -C oracle ref` together with the reference sources where
// they lie (src/PA/SPA/SerialCopyModel.cpp, SerialCopyModel3.cpp,
// src/Common/Timer.cpp). It mirrors how src/main.cu:93-98 calls the serial
// copy model: construct SerialCopyModel(n, m, p) and call generateGraph(conf)
// with the degree-distribution flag set, so the reference writes its own
// "Degree,Count" CSV (PAHelper.hpp:15-67). Its draws come from MT19937-64 with
// the default seed (Random64.hpp:75-76), so its output is compared with ours
// statistically, not bit for bit.
//
// usage: scm_ref <n> <m> <p> <outdir/> [write_dd=1]
#include <cstdlib>
#include <iostream>

#include "PA/SPA/SerialCopyModel.h"

int main(int argc, char **argv)
{
    if (argc != 5 && argc != 6) {
        std::cerr << "usage: scm_ref <n> <m> <p> <outdir/> [write_dd=1]" << std::endl;
        return 2;
    }
    ExperimentConfiguration conf;
    conf._n1 = (NodeType)std::atoll(argv[1]);
    conf._x1 = (WeightType)std::atoi(argv[2]);
    conf._alpha = std::atof(argv[3]);
    conf._outputDir = argv[4];
    conf._distfile1 = "scm_dd.csv";
    conf._FLAG_SHOW_DEGREE_DISTRIBUTION = argc == 6 ? std::atoi(argv[5]) != 0 : true;
    conf._N_STAGE_TIMINGS = 0;
    SerialCopyModel scm(conf._n1, (EdgeIndexType)conf._x1, conf._alpha);
    scm.generateGraph(conf);
    return 0;
}
