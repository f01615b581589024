// RUN_MINIMIZER_SMOKE of the reference's tests/run_test/run_test.cpp:98-152:
// two type-0 particles at (1,1) and (3,1) in a 10x10 box, FIRE from
// FireParams::from_config, force tolerance 1e-3.  argv[1]: max steps.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "md_energy_minimizer.hpp"

int main(int argc, char **argv) {
    const int max_steps = argc > 1 ? std::atoi(argv[1]) : 200;
    MDConfig mini_cfg{};
    mini_cfg.rank_idx = 0;
    mini_cfg.rank_size = 1;
    mini_cfg.box_w_global = 10.0;
    mini_cfg.box_h_global = 10.0;
    mini_cfg.dt = 1e-3;
    mini_cfg.cutoff = 2.5;
    mini_cfg.EPSILON_AB = 1.0;
    mini_cfg.THREADS_PER_BLOCK = 128;
    mini_cfg.n_particles_global = 2;
    mini_cfg.n_local = 2;
    mini_cfg.n_cap = 16;
    mini_cfg.halo_left_cap = 8;
    mini_cfg.halo_right_cap = 8;
    mini_cfg.x_min = 0.0;
    mini_cfg.x_max = mini_cfg.box_w_global;
    MDConfigManager mini_mgr(mini_cfg);
    EnergyMinimizer minimizer(mini_mgr, 0);
    std::vector<Particle> frame(2);
    frame[0] = Particle{};
    frame[1] = Particle{};
    frame[0].pos = {1.0, 1.0};
    frame[1].pos = {3.0, 1.0};
    frame[0].type = frame[1].type = 0;
    frame[0].vel = frame[1].vel = make_double2(0.0, 0.0);
    frame[0].acc = frame[1].acc = make_double2(0.0, 0.0);
    FireParams params = FireParams::from_config(mini_cfg);
    auto result = minimizer.minimize_frame(frame, 1e-3, max_steps, params);
    const double r = std::fabs(result.frame[1].pos.x - result.frame[0].pos.x);
    std::printf("[MinimizerSmoke] steps=%d max_force=%.3e converged=%d r=%.8f U=%.10f samples=%zu\n", result.steps,
                result.max_force, result.converged ? 1 : 0, r, result.potential_energy, result.energy_steps.size());
    bool threw = false;
    try {
        minimizer.minimize_frame(frame, 1e-3, 0, params);
    } catch (const std::exception &) {
        threw = true;
    }
    return threw ? 0 : 1;
}
