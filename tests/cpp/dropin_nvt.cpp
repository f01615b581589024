// The reference's driver loop on this library: MDSimulation constructed from a
// config (init_particles start), then one step_single_nose_hoover() per loop
// iteration, as run/series_partial_U_lambda_test.cpp:207-219 steps, with the
// record-interval getters (cal_total_U, cal_total_K, cal_partial_U_lambda)
// every `record` steps (0: none).  Prints one JSON line with the timed rate.
// argv: n_particles warmup steps [record]
// Workload: 2:1 periodic box at density 0.75, N/2 type-0 particles in the
// central slab (devide_p 0.5), T_init = T_target = 0.6, Q = 10, dt = 0.005,
// eps_AB = 0.5, cutoff 2.5.  The pair-pass tuning follows MDX_TUNING (unset:
// autotuned at the first step, before the timed loop).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "md_env.hpp"

int main(int argc, char **argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: dropin_nvt n_particles warmup steps [record]\n");
        return 1;
    }
    const int n = std::atoi(argv[1]);
    const int warmup = std::atoi(argv[2]);
    const int steps = std::atoi(argv[3]);
    const int record = argc > 4 ? std::atoi(argv[4]) : 0;
    MDConfigManager cfg_mgr;
    MDConfig &c = cfg_mgr.config;
    c.n_particles_global = n;
    c.n_particles_type0 = n / 2;
    c.box_h_global = std::sqrt(n / 0.75 / 2.0);
    c.box_w_global = 2.0 * c.box_h_global;
    c.T_init = 0.6;
    c.T_target = 0.6;
    c.EPSILON_AB = 0.5;
    c.dt = 0.005;
    c.Q = 10.0;
    c.cutoff = 2.5;
    c.devide_p = 0.5;
    c.rank_size = 1;
    c.rank_idx = 0;
    MDSimulation sim(cfg_mgr, 0 /* communicator placeholder */);
    double sink = 0.0;
    for (int step = 0; step < warmup; ++step) sim.step_single_nose_hoover();
    sink += sim.cal_total_K();   // synchronises: the warm-up (and tuning) is done
    const auto t0 = std::chrono::steady_clock::now();
    for (int step = 0; step < steps; ++step) {
        sim.step_single_nose_hoover();
        if (record > 0 && step % record == 0) {
            sink += sim.cal_total_U() + sim.cal_total_K() + sim.cal_partial_U_lambda(0.1);
        }
    }
    const double K = sim.cal_total_K();   // synchronises
    const auto t1 = std::chrono::steady_clock::now();
    const double wall = std::chrono::duration<double>(t1 - t0).count();
    mdx_tuning tu{};
    mdx_get_tuning(sim.handle(), &tu);
    std::printf("{\"n\": %d, \"steps\": %d, \"warmup\": %d, \"record\": %d, \"wall_s\": %.6f, "
                "\"ms_per_step\": %.6f, \"atom_steps_per_s\": %.6e, \"K_per_atom\": %.6f, \"finite\": %s, "
                "\"tuning\": {\"cell_scale\": %g, \"lanes_per_particle\": %d, \"skin\": %g, \"window\": %d}}\n",
                n, steps, warmup, record, wall, wall * 1e3 / steps, static_cast<double>(n) * steps / wall, K / n,
                std::isfinite(sink + K) ? "true" : "false", tu.cell_scale, tu.lanes_per_particle, tu.skin,
                tu.window);
    return std::isfinite(sink + K) ? 0 : 2;
}
