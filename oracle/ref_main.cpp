// ref_main.cpp -- entry point for oracle/_ref/dCNS_ref, a build of the
// reference's command functions (controlLayer4.cpp) used only to produce
// golden outputs (SAM / o files, ranSco scores) for the tests.
// Dispatches like the reference's main.cpp:8-50, minus the gtest runner.
#include <cstring>
#include <iostream>
#include "controlLayer.h"

int main(int argc, char **argv) {
    if (argc <= 1) { std::cerr << "usage: dCNS_ref <cut1Gap|cut2Gaps|pairCns2Gaps|pairCnsXExtend|ranSco|weighted1Gap|weighted2Gaps|cut2Gaps2|maskGenome> ...\n"; return 1; }
    const char *p = argv[1];
    if (!std::strcmp(p, "cut1Gap")) return cut1Gap(argc - 1, argv + 1);
    if (!std::strcmp(p, "cut2Gaps")) return cut2Gaps(argc - 1, argv + 1);
    if (!std::strcmp(p, "pairCns2Gaps")) return pairCns2Gaps(argc - 1, argv + 1);
    if (!std::strcmp(p, "pairCnsXExtend")) return pairCnsXExtend(argc - 1, argv + 1);
    if (!std::strcmp(p, "ranSco")) return smitherWaterManScoreOfRandomFragments(argc - 1, argv + 1);
    if (!std::strcmp(p, "weighted1Gap")) return weighted1Gap(argc - 1, argv + 1);
    if (!std::strcmp(p, "cut2Gaps2")) return cut2Gaps2(argc - 1, argv + 1);
    if (!std::strcmp(p, "weighted2Gaps")) return weighted2Gaps(argc - 1, argv + 1);
    if (!std::strcmp(p, "maskGenome")) return maskGenome(argc - 1, argv + 1);
    std::cerr << "unknown command " << p << "\n";
    return 1;
}
