// EnergyMinimizer (FIRE) of the reference (include/md_energy_minimizer.hpp:11-97
// there), running on one GPU through the mdx C ABI.
#pragma once

#include <string>
#include <vector>

#include "md_config.hpp"
#include "md_particle.hpp"

struct FireParams {
    double dt_init = 0.0;
    double dt_max = 0.0;
    double alpha0 = 0.1;
    double f_inc = 1.1;
    double f_dec = 0.5;
    int N_min = 5;
    // Overlap policy of this library (0 keeps the reference behaviour).
    double max_step = 0.0;      // cap on |dt v| per FIRE step
    double fp64_switch = 0.0;   // evaluate pairs in FP64 once max|a| falls below this

    static FireParams from_config(const MDConfig &cfg) {
        FireParams params;
        params.dt_init = cfg.dt;
        params.dt_max = 5.0 * cfg.dt;
        return params;
    }
};

struct MinimizeResult {
    std::vector<Particle> frame;
    std::vector<int> energy_steps;
    std::vector<double> energy_trace;
    int steps = 0;
    bool converged = false;
    double max_force = 0.0;
    double potential_energy = 0.0;
};

class EnergyMinimizer {
public:
    explicit EnergyMinimizer(MDConfigManager config_manager);
    template <class Comm>
    EnergyMinimizer(MDConfigManager config_manager, Comm) : EnergyMinimizer(std::move(config_manager)) {}

    MinimizeResult minimize_frame(const std::vector<Particle> &frame, double force_tol, int max_steps,
                                  const FireParams &params);
    void sample_collect(std::vector<Particle> &host_out, bool gather_root = true);

    void set_precision(int precision) { precision_ = precision; }

private:
    MDConfigManager cfg_manager;
    std::vector<Particle> last_frame_;
    int precision_ = 1;   // MDX_PREC_ACC64
};
