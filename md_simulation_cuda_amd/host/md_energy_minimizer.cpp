// EnergyMinimizer::minimize_frame (src/md_energy_minimizer.cu:122-272 of the
// reference) on one GPU through mdx_fire.  Errors the reference throws are
// thrown here with the same messages.
#include "md_energy_minimizer.hpp"

#include <stdexcept>
#include <string>

#include "mdx.h"

EnergyMinimizer::EnergyMinimizer(MDConfigManager config_manager) : cfg_manager(std::move(config_manager)) {
    if (cfg_manager.config.n_cap <= 0) throw std::runtime_error("EnergyMinimizer: n_cap must be positive.");
}

MinimizeResult EnergyMinimizer::minimize_frame(const std::vector<Particle> &frame, double force_tol,
                                               int max_steps, const FireParams &params) {
    if (max_steps <= 0) throw std::runtime_error("EnergyMinimizer: max_steps must be positive.");
    if (static_cast<long long>(frame.size()) > cfg_manager.config.n_cap)
        throw std::runtime_error("EnergyMinimizer: input frame exceeds buffer capacity.");
    MinimizeResult result;
    if (frame.empty()) {   // no particles: max|a| = 0 converges at step 0
        result.energy_steps = {0};
        result.energy_trace = {0.0};
        if (0.0 < force_tol) {
            result.converged = true;
            result.energy_steps.push_back(1);
            result.energy_trace.push_back(0.0);
        } else {
            result.steps = max_steps;
            result.energy_steps.push_back(max_steps);
            result.energy_trace.push_back(0.0);
        }
        last_frame_.clear();
        return result;
    }
    const MDConfig &c = cfg_manager.config;
    mdx_params p{};
    p.Lx = c.box_w_global;
    p.Ly = c.box_h_global;
    p.sigma_aa = c.SIGMA_AA;
    p.sigma_bb = c.SIGMA_BB;
    p.sigma_ab = c.SIGMA_AB;
    p.eps_aa = c.EPSILON_AA;
    p.eps_bb = c.EPSILON_BB;
    p.eps_ab = c.EPSILON_AB;
    p.mass_a = c.MASS_A;
    p.mass_b = c.MASS_B;
    p.cutoff = c.cutoff;
    p.dt = c.dt;
    p.T_target = c.T_target;
    p.Q = c.Q;
    p.barostat_mass = c.barostat_mass;
    p.k_piston = c.k_piston;
    p.barostat_height_min_ratio = c.barostat_height_min_ratio;
    p.barostat_height_max_ratio = c.barostat_height_max_ratio;
    p.devide_p = c.devide_p;
    p.cell_scale = 1.0;
    p.precision = precision_;
    p.device = -1;
    mdx_sim *sim = nullptr;
    int rc = mdx_create(&p, reinterpret_cast<const mdx_particle64 *>(frame.data()),
                        static_cast<int64_t>(frame.size()), &sim);
    if (rc != MDX_OK) throw std::runtime_error(std::string("EnergyMinimizer: ") + mdx_last_error());
    mdx_fire_params fp{};
    fp.dt_init = params.dt_init;
    fp.dt_max = params.dt_max;
    fp.alpha0 = params.alpha0;
    fp.f_inc = params.f_inc;
    fp.f_dec = params.f_dec;
    fp.N_min = params.N_min;
    fp.energy_interval = 1000;
    fp.max_step = params.max_step;
    fp.fp64_switch = params.fp64_switch;
    const int cap = max_steps / 1000 + 4;
    std::vector<int32_t> es(static_cast<size_t>(cap));
    std::vector<double> et(static_cast<size_t>(cap));
    mdx_fire_result fr{};
    rc = mdx_fire(sim, &fp, force_tol, max_steps, &fr, es.data(), et.data(), cap);
    if (rc == MDX_OK) {
        result.frame.resize(frame.size());
        rc = mdx_collect(sim, reinterpret_cast<mdx_particle64 *>(result.frame.data()),
                         static_cast<int64_t>(frame.size()));
    }
    const std::string err = rc == MDX_OK ? "" : mdx_last_error();
    mdx_destroy(sim);
    if (rc != MDX_OK) throw std::runtime_error("EnergyMinimizer: " + err);
    const int k = fr.n_energy < cap ? fr.n_energy : cap;
    result.energy_steps.assign(es.begin(), es.begin() + k);
    result.energy_trace.assign(et.begin(), et.begin() + k);
    result.steps = fr.steps;
    result.converged = fr.converged != 0;
    result.max_force = fr.max_force;
    result.potential_energy = fr.potential_energy;
    last_frame_ = result.frame;
    return result;
}

void EnergyMinimizer::sample_collect(std::vector<Particle> &host_out, bool) { host_out = last_frame_; }
