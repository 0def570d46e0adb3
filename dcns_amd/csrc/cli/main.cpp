// main.cpp -- the dCNS command-line drop-in for the commands this library
// implements on the GPU (dispatch as in the reference's main.cpp:8-50).
#include <cstring>
#include <iostream>

#include "../../../include/dcns.h"

static void usage() {
    std::cout << "Program: dCNS (MI355X build, " << dcns_version() << ")\n"
              << "Usage: dCNS <command> [options]\n"
              << "Commands:\n"
              << "    cut1Gap             masking genome, sliding window seed, x-drop extension\n"
              << "    ranSco              smith-waterman scores of random sequence fragments\n"
              << "                        for lambda and k fitting\n";
}

int main(int argc, char **argv) {
    if (argc <= 1) {
        usage();
        return 1;
    }
    const char *p = argv[1];
    if (!std::strcmp(p, "-h") || !std::strcmp(p, "--help")) {
        usage();
        return 1;
    }
    if (!std::strcmp(p, "cut1Gap")) return dcns_cut1gap_main(argc - 1, argv + 1);
    if (!std::strcmp(p, "ranSco")) return dcns_ransco_main(argc - 1, argv + 1);
    std::cerr << "dCNS: command " << p << " is not part of this build" << std::endl;
    usage();
    return 1;
}
