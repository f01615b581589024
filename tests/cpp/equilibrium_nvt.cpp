// check_eqlibrium through the header-compatible MDSimulation, driven as the
// reference's run_test does (tests/run_test/run_test.cpp:207-216): NVT steps
// one call at a time, the test every `every` steps.  One species, N = 1024 at
// density 0.7, T = 0.7, dt = 0.005, save_dt_interval = 1 (a window is 100
// samples, 20000 steps).  Prints the step of the first positive test.
// argv: max_steps every
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "md_env.hpp"

int main(int argc, char **argv) {
    if (argc != 3) return 2;
    const int max_steps = std::atoi(argv[1]), every = std::atoi(argv[2]);
    MDConfigManager m;
    MDConfig &c = m.config;
    c.n_particles_global = 1024;
    c.n_particles_type0 = 1024;
    c.box_w_global = std::sqrt(1024 / 0.7);
    c.box_h_global = c.box_w_global;
    c.T_init = 0.7;
    c.T_target = 0.7;
    c.dt = 0.005;
    c.Q = 10.0;
    c.devide_p = 1.0;
    c.save_dt_interval = 1.0;
    c.rank_size = 1;
    c.rank_idx = 0;
    MDSimulation sim(m, 0);
    int first = -1, checks = 0, positives = 0;
    for (int step = 1; step <= max_steps; ++step) {
        sim.step_single_nose_hoover();
        if (step % every == 0) {
            ++checks;
            const bool eq = sim.check_eqlibrium(1.0);
            positives += eq;
            if (eq && first < 0) first = step;
        }
    }
    std::printf("equilibrium first=%d checks=%d positives=%d t=%.3f U=%.6f\n", first, checks, positives, sim.t,
                sim.cal_total_U() / 1024);
    return first > 0 ? 0 : 1;
}
