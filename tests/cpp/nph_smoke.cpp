// The reference's tests/nph/nph_smoke.cpp and tests/nph_piston/nph_smoke_piston.cpp
// on this library: 100 Nose-Hoover steps averaging the instantaneous
// pressure, then 200 NPH (or piston) steps at that pressure.  Exit codes 2/3
// on non-finite state, as there.  argv: config.json [piston]
#include <cmath>
#include <cstdio>
#include <string>

#include "md_env.hpp"

int main(int argc, char **argv) {
    if (argc < 2) return 1;
    const bool piston = argc > 2 && std::string(argv[2]) == "piston";
    MDConfigManager cfg_mgr = MDConfigManager::config_from_json(argv[1]);
    cfg_mgr.config.rank_size = 1;
    cfg_mgr.config.rank_idx = 0;
    MDSimulation sim(cfg_mgr, 0 /* communicator placeholder */);
    const int auto_steps = 100, production_steps = 200;
    double target_pressure = 0.0;
    for (int step = 0; step < auto_steps; ++step) {
        sim.step_single_nose_hoover();
        const double p_inst = sim.cal_instant_pressure();
        target_pressure += p_inst;
        if (!std::isfinite(p_inst)) {
            std::fprintf(stderr, "[nph_smoke] non-finite pressure during AUTO_TARGET at step %d\n", step);
            return 2;
        }
    }
    target_pressure /= static_cast<double>(auto_steps);
    for (int step = 0; step < production_steps; ++step) {
        if (piston)
            sim.step_single_NPH_piston(target_pressure);
        else
            sim.step_single_NPH(target_pressure);
        const double p_inst = sim.get_last_instant_pressure();
        if (!std::isfinite(p_inst) || !std::isfinite(sim.get_Lx()) || !std::isfinite(sim.get_Ly()) ||
            sim.get_Lx() <= 0.0 || sim.get_Ly() <= 0.0 ||
            (piston && (!std::isfinite(sim.get_nph_piston_velocity()) ||
                        !std::isfinite(sim.get_nph_piston_force())))) {
            std::fprintf(stderr, "[nph_smoke] invalid state at step %d: Lx=%g Ly=%g P=%g\n", step, sim.get_Lx(),
                         sim.get_Ly(), p_inst);
            return 3;
        }
    }
    sim.sample_collect();
    std::vector<Particle> frame;
    sim.get_host_particles(frame);
    std::printf("[nph_smoke] completed. Lx=%.10g, Ly=%.10g, P_target=%.10g, P_last=%.10g, dLy_dt=%.6g, "
                "F_piston=%.6g, t=%.6g, n=%zu, K=%.10g, U=%.10g\n",
                sim.get_Lx(), sim.get_Ly(), target_pressure, sim.get_last_instant_pressure(),
                sim.get_nph_piston_velocity(), sim.get_nph_piston_force(), sim.t, frame.size(), sim.cal_total_K(),
                sim.cal_total_U());
    return 0;
}
