// EquilibriumTracker on a given series (CPU): equilibrium_stats interval dt
// sensitivity n_particles series.txt -> one line per sample after recording
// it and running one test: "<index> <streak> <0|1>"; the first line gives the
// window length.
#include <cstdio>
#include <cstdlib>
#include <fstream>

#include "md_equilibrium.hpp"

int main(int argc, char **argv) {
    if (argc != 6) return 2;
    EquilibriumTracker tr;
    tr.reset(std::atof(argv[1]), std::atof(argv[2]));
    const double sens = std::atof(argv[3]);
    const int n = std::atoi(argv[4]);
    std::ifstream in(argv[5]);
    std::printf("window %zu\n", tr.window());
    double U;
    for (int i = 0; in >> U; ++i) {
        tr.record(U);
        const bool eq = tr.evaluate(sens, n);
        std::printf("%d %d %d\n", i, tr.streak(), eq ? 1 : 0);
    }
    return 0;
}
