// get_N_profile, get_density_profile and get_pressure_profile through the
// header-compatible MDSimulation (reference src/md_env.cu:3920-3988,
// 4086-4104, 3821-3918): 20 NVE steps, then 16-slab profiles; the counts
// cover every particle, the density profile is (n_A - n_B) / (n_A + n_B) per
// slab (0 where empty), so it lies in [-1, 1], and the pressure profile's
// slabs sum to the scalar pressure.  argv: config.json
#include <cmath>
#include <cstdio>
#include <vector>

#include "md_env.hpp"

int main(int argc, char **argv) {
    if (argc < 2) return 1;
    MDConfigManager cfg_mgr = MDConfigManager::config_from_json(argv[1]);
    cfg_mgr.config.rank_size = 1;
    cfg_mgr.config.rank_idx = 0;
    MDSimulation sim(cfg_mgr, 0 /* communicator placeholder */);
    for (int s = 0; s < 20; ++s) sim.step_single_NVE();
    const int nb = 16;
    const std::vector<int> n = sim.get_N_profile(nb);
    const std::vector<double> d = sim.get_density_profile(nb);
    if (static_cast<int>(n.size()) != 2 * nb || static_cast<int>(d.size()) != nb) return 2;
    long total = 0, total_a = 0;
    for (int i = 0; i < nb; ++i) {
        total += n[i] + n[i + nb];
        total_a += n[i];
        const int a = n[i], b = n[i + nb];
        const double want = a + b == 0 ? 0.0 : static_cast<double>(a - b) / static_cast<double>(a + b);
        if (!(std::fabs(d[i]) <= 1.0) || d[i] != want) {
            std::fprintf(stderr, "slab %d: density %g, counts %d %d\n", i, d[i], a, b);
            return 3;
        }
    }
    if (total != cfg_mgr.config.n_particles_global) return 4;
    const std::vector<double> pp = sim.get_pressure_profile(nb);
    if (static_cast<int>(pp.size()) != 3 * nb || !sim.get_pressure_profile(0).empty()) return 5;
    double tr = 0.0;
    for (int i = 0; i < nb; ++i) tr += pp[i] + pp[i + nb];
    const double P_from_profile = tr / nb / 2.0, P = sim.cal_instant_pressure();
    if (!(std::fabs(P_from_profile - P) <= 1e-6 * std::fabs(P))) {
        std::fprintf(stderr, "pressure profile sums to %.12g, scalar pressure %.12g\n", P_from_profile, P);
        return 6;
    }
    std::printf("profile ok n=%ld a=%ld slab0=%d,%d P=%.9g\n", total, total_a, n[0], n[nb], P);
    return 0;
}
