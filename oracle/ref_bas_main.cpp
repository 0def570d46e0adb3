// ref_bas_main.cpp -- runs the reference's SmithWaterman_bas body
// (impl/SmithWaterman_bas.cpp:59-201, compiled where it lies by
// oracle/Makefile into oracle/_ref/swbas_ref) on one case read from stdin:
//
//   L1 L2 open ext reverse
//   <L1 codes>
//   <L2 codes>
//
// and prints "RESULT maxScore end1 end2 ncigar cigar..." after the
// function's own output.  Test infrastructure only: one case per process, so
// that a case whose traceback never ends can be stopped by a time limit.
#include <cstdint>
#include <iostream>
#include <vector>

// The body's signature (the header declares an 11-argument variant with no body).
std::vector<uint32_t> SmithWaterman(int8_t *seq1, int8_t *seq2, const uint32_t &length1, const uint32_t &length2,
                                    const int &_open_gap_penalty, const int &_extend_gap_penalty, uint32_t &maxScore,
                                    uint32_t &endPosition1, uint32_t &endPosition2, bool reverseAlignment);

int main() {
    uint32_t l1 = 0, l2 = 0;
    int open = 0, ext = 0, rev = 0;
    if (!(std::cin >> l1 >> l2 >> open >> ext >> rev)) return 2;
    std::vector<int8_t> a(l1 + 1), b(l2 + 1);
    for (uint32_t i = 0; i < l1; ++i) {
        int v;
        std::cin >> v;
        a[i] = (int8_t)v;
    }
    for (uint32_t i = 0; i < l2; ++i) {
        int v;
        std::cin >> v;
        b[i] = (int8_t)v;
    }
    uint32_t mx = 0, e1 = 0, e2 = 0;
    const std::vector<uint32_t> cig = SmithWaterman(a.data(), b.data(), l1, l2, open, ext, mx, e1, e2, rev != 0);
    std::cout << "RESULT " << mx << " " << e1 << " " << e2 << " " << cig.size();
    for (uint32_t c : cig) std::cout << " " << c;
    std::cout << std::endl;
    return 0;
}
