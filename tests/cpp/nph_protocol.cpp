// The three-phase NPH protocol (NphProtocol) on a config file, with the
// phase lengths shortened from the command line for a test run:
// nph_protocol config.json out_dir auto_steps eval_every window tol_rel
//              stable_needed warm_min warm_max nph_steps save_dt_interval
// Prints the phase switch steps and the final box.
#include <cstdio>
#include <cstdlib>

#include "md_nph_protocol.hpp"

int main(int argc, char **argv) {
    if (argc != 12) {
        std::fprintf(stderr, "usage: nph_protocol config.json out_dir auto eval window tol stable warm_min "
                             "warm_max nph_steps save_dt_interval\n");
        return 2;
    }
    MDConfigManager m = MDConfigManager::config_from_json(argv[1]);
    MDConfig &c = m.config;
    c.rank_size = 1;
    c.rank_idx = 0;
    c.pressure_target_auto_steps = std::atoi(argv[3]);
    c.pressure_eval_interval_steps = std::atoi(argv[4]);
    c.pressure_window_samples = std::atoi(argv[5]);
    c.pressure_tolerance_rel = std::atof(argv[6]);
    c.pressure_stable_required_windows = std::atoi(argv[7]);
    c.nph_warmup_min_steps = std::atoi(argv[8]);
    c.nph_warmup_max_steps = std::atoi(argv[9]);
    c.nph_steps = std::atoi(argv[10]);
    c.save_dt_interval = std::atof(argv[11]);
    MDSimulation sim(m, 0);
    NphProtocol prot(sim, c, argv[2]);
    const long long n = prot.run();
    using P = NphProtocol::Phase;
    std::printf("nph_protocol steps=%lld finished=%d P_target=%.10e warmup_at=%lld production_at=%lld forced=%d "
                "auto=%lld warm=%lld nph=%lld Lx=%.10e Ly=%.10e P_last=%.10e\n",
                n, prot.finished() ? 1 : 0, prot.P_target(), static_cast<long long>(prot.switch_step(P::NPH_WARMUP)),
                static_cast<long long>(prot.switch_step(P::NPH_PRODUCTION)), prot.warmup_forced() ? 1 : 0,
                static_cast<long long>(prot.steps_in(P::AUTO_TARGET)),
                static_cast<long long>(prot.steps_in(P::NPH_WARMUP)),
                static_cast<long long>(prot.steps_in(P::NPH_PRODUCTION)), sim.get_Lx(), sim.get_Ly(),
                sim.get_last_instant_pressure());
    return prot.finished() ? 0 : 1;
}
