// get_N_profile and get_density_profile through the header-compatible
// MDSimulation (reference src/md_env.cu:3920-3988, 4086-4104): 20 NVE steps,
// then 16-slab profiles; the counts cover every particle and the density
// profile is (n_A - n_B) / (n_A + n_B) per slab (0 where empty), so it lies
// in [-1, 1].  argv: config.json
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
    std::printf("profile ok n=%ld a=%ld slab0=%d,%d\n", total, total_a, n[0], n[nb]);
    return 0;
}
