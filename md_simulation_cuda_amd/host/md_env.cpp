// MDSimulation on one GPU through the mdx C ABI.
//
// Constructor semantics follow src/md_env.cu:21-164 (init_particles, resume
// from a PFRAME file), the steps src/md_env.cu:1576-2023 and the getters
// src/md_env.cu:2141-2514.  Where the reference calls MPI_Abort or exits on a
// CUDA error, this class prints the message and exits with status 1.
#include "md_env.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>

namespace {

[[noreturn]] void abort_run(const std::string &msg) {
    std::fprintf(stderr, "%s\n", msg.c_str());
    std::fflush(stderr);
    std::exit(1);
}

int precision_from_env() {
    const char *e = std::getenv("MDX_PRECISION");
    if (!e) return MDX_PREC_COMP32;
    const std::string s(e);
    if (s == "acc64" || s == "1") return MDX_PREC_ACC64;
    if (s == "fp64" || s == "2") return MDX_PREC_FP64;
    return MDX_PREC_COMP32;
}

mdx_params params_from(const MDConfig &c, int precision) {
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
    p.barostat_area_rate_init = c.barostat_area_rate_init;
    p.barostat_height_min_ratio = c.barostat_height_min_ratio;
    p.barostat_height_max_ratio = c.barostat_height_max_ratio;
    p.devide_p = c.devide_p;
    p.cell_scale = 1.0;
    p.precision = precision;
    p.device = -1;
    return p;
}

}  // namespace

void MDSimulation::check(int rc, const char *where, bool throws) const {
    if (rc == MDX_OK) return;
    const std::string msg = std::string("[") + where + "] " + mdx_last_error();
    if (throws) throw std::runtime_error(msg);
    abort_run(msg);
}

void MDSimulation::create_from_frame(const std::vector<Particle> &frame) {
    // equilibrium record (src/md_env.cu:174-195): time 0, first record at one interval
    t = 0.0;
    eq_.reset(cfg_manager.config.save_dt_interval, cfg_manager.config.dt);
    next_record_time_ = eq_.interval();
    const mdx_params p = params_from(cfg_manager.config, precision_);
    check(mdx_create(&p, reinterpret_cast<const mdx_particle64 *>(frame.data()),
                     static_cast<int64_t>(frame.size()), &sim_),
          "MDSimulation");
    // tuning of the pair pass: MDX_TUNING unset or "auto" (autotune at the
    // first NVE / NVT step), "off", or "cell_scale,lanes,skin,window"
    const char *tv = std::getenv("MDX_TUNING");
    const std::string spec = tv ? tv : "auto";
    if (spec == "auto" || spec.empty()) {
        autotune_pending_ = true;
    } else if (spec != "off") {
        double cs = 1.0, skin = 0.0;
        int lanes = 1, window = 0;
        if (std::sscanf(spec.c_str(), "%lf,%d,%lf,%d", &cs, &lanes, &skin, &window) < 3)
            abort_run("[MDSimulation] MDX_TUNING must be auto, off or cell_scale,lanes,skin[,window]: " + spec);
        set_tuning(cs, lanes, skin, window);
    }
    MDConfig &c = cfg_manager.config;
    c.rank_idx = 0;
    c.n_local = static_cast<int>(frame.size());
    c.n_halo_left = 0;
    c.n_halo_right = 0;
    c.x_min = 0.0;
    c.x_max = c.box_w_global;
}

MDSimulation::MDSimulation(MDConfigManager config_manager) : cfg_manager(std::move(config_manager)) {
    precision_ = precision_from_env();
    const MDConfig &c = cfg_manager.config;
    if (c.box_w_global <= 0.0) abort_run("[broadcast_params] Error: box_w_global <= 0.");
    if (c.n_particles_global <= 0 || c.n_particles_type0 < 0 || c.n_particles_type0 > c.n_particles_global)
        abort_run("[init_particles] invalid particle counts");
    h_particles.assign(static_cast<size_t>(c.n_particles_global), Particle{});
    mdx_init_params ip{};
    ip.n_particles_global = c.n_particles_global;
    ip.n_particles_type0 = c.n_particles_type0;
    ip.box_w = c.box_w_global;
    ip.box_h = c.box_h_global;
    ip.T_init = c.T_init;
    ip.mass_a = c.MASS_A;
    ip.mass_b = c.MASS_B;
    ip.devide_p = c.devide_p;
    check(mdx_init_particles(&ip, reinterpret_cast<mdx_particle64 *>(h_particles.data())), "init_particles");
    create_from_frame(h_particles);
}

MDSimulation::MDSimulation(MDConfigManager config_manager, const std::string &filename, int step)
    : cfg_manager(std::move(config_manager)) {
    precision_ = precision_from_env();
    const MDConfig &c = cfg_manager.config;
    if (c.box_w_global <= 0.0) abort_run("[broadcast_params] Error: box_w_global <= 0.");
    std::vector<Particle> loaded;
    bool success = false;
    try {   // src/md_env.cu:103-130
        FileReader reader(filename);
        if (!reader.corrupted()) {
            std::vector<Particle> buf;
            std::uint64_t idx = 0;
            while (reader.next_frame(buf, idx)) {
                if (static_cast<int>(idx) == step) {
                    loaded = std::move(buf);
                    success = true;
                    break;
                }
            }
            if (!success)
                std::fprintf(stderr, "Error: Frame index %d not found in %s.\n", step, filename.c_str());
        } else {
            std::fprintf(stderr, "Error: File %s is corrupted.\n", filename.c_str());
        }
    } catch (const std::exception &e) {
        std::fprintf(stderr, "Error opening file %s: %s\n", filename.c_str(), e.what());
        success = false;
    }
    if (!success) abort_run("[MDSimulation] resume failed");
    if (loaded.size() != static_cast<size_t>(c.n_particles_global))
        abort_run("[Error] Loaded particle count " + std::to_string(loaded.size()) +
                  " doesn't match n_particles_global " + std::to_string(c.n_particles_global) + ".");
    h_particles = loaded;
    create_from_frame(h_particles);
}

MDSimulation::~MDSimulation() {
    if (sim_) mdx_destroy(sim_);
}

void MDSimulation::sync_scalars() {
    mdx_obs o{};
    check(mdx_observables(sim_, 0u, &o), "observables");
    last_instant_pressure = o.last_P;
    nph_area_momentum = o.p_A;
    nph_piston_velocity = o.v_p;
    nph_piston_force = o.F_p;
    cfg_manager.config.box_w_global = o.Lx;
    cfg_manager.config.box_h_global = o.Ly;
    cfg_manager.config.x_max = o.Lx;
}

void MDSimulation::set_tuning(double cell_scale, int lanes_per_particle, double skin, int window) {
    mdx_tuning tu{};
    tu.cell_scale = cell_scale;
    tu.lanes_per_particle = lanes_per_particle;
    tu.skin = skin;
    tu.window = window;
    check(mdx_set_tuning(sim_, &tu), "set_tuning", true);
    autotune_pending_ = false;
}

void MDSimulation::autotune(int kind, int steps_per_candidate) {
    mdx_tune_report rep{};
    check(mdx_autotune(sim_, kind, steps_per_candidate, &rep), "autotune", true);
    autotune_pending_ = false;
}

void MDSimulation::middle_reflect_LG() { check(mdx_middle_reflect(sim_), "middle_reflect_LG"); }

void MDSimulation::step_many(int kind, int nsteps, double pressure_ext, bool do_middle_wrap) {
    if (nsteps <= 0) return;
    if (autotune_pending_ && (kind == MDX_NVE || kind == MDX_NVT)) autotune(kind);
    check(mdx_step(sim_, kind, nsteps, pressure_ext, do_middle_wrap ? MDX_STEP_MIDDLE_WRAP : 0u), "step");
    t += cfg_manager.config.dt * nsteps;
    if (kind == MDX_NPH || kind == MDX_PISTON) sync_scalars();
}

void MDSimulation::step_single() { step_single_NVE(); }
void MDSimulation::step_single_NVE() { step_many(MDX_NVE, 1); }
void MDSimulation::step_single_nose_hoover(bool do_middle_wrap) { step_many(MDX_NVT, 1, 0.0, do_middle_wrap); }
void MDSimulation::step_single_NPH(double pressure_ext) { step_many(MDX_NPH, 1, pressure_ext); }
void MDSimulation::step_single_NPH_piston(double pressure_ext) { step_many(MDX_PISTON, 1, pressure_ext); }

double MDSimulation::get_nph_area_rate() const {
    const double W = cfg_manager.config.barostat_mass > 1e-12 ? cfg_manager.config.barostat_mass : 1e-12;
    return nph_area_momentum / W;
}

double MDSimulation::cal_total_K() {
    mdx_obs o{};
    check(mdx_observables(sim_, MDX_OBS_K, &o), "cal_total_K");
    return o.K;
}

double MDSimulation::cal_total_U() {
    mdx_obs o{};
    check(mdx_observables(sim_, MDX_OBS_U, &o), "cal_total_U");
    return o.U;
}

double MDSimulation::cal_instant_pressure() {
    mdx_obs o{};
    check(mdx_observables(sim_, MDX_OBS_P, &o), "cal_instant_pressure");
    last_instant_pressure = o.P;
    nph_piston_force = o.F_p;
    return o.P;
}

double MDSimulation::cal_partial_U_lambda(double epsilon_lambda) {
    double v = 0.0;
    check(mdx_partial_U_lambda(sim_, epsilon_lambda, &v), "cal_partial_U_lambda");
    return v;
}

std::vector<int> MDSimulation::get_N_profile(int n_bins_per_rank) {
    std::vector<int> out(static_cast<size_t>(n_bins_per_rank > 0 ? 2 * n_bins_per_rank : 0));
    if (n_bins_per_rank > 0)
        check(mdx_N_profile(sim_, n_bins_per_rank, reinterpret_cast<int32_t *>(out.data())), "get_N_profile");
    return out;
}

bool MDSimulation::check_eqlibrium(double sensitivity) {   // src/md_env.cu:2478-2500
    const double sens = std::max(sensitivity, 1e-3);
    while (t + 1e-12 >= next_record_time_) {   // the same U for every boundary passed
        eq_.record(cal_total_U());
        next_record_time_ += eq_.interval();
    }
    return eq_.evaluate(sens, cfg_manager.config.n_particles_global);
}

std::vector<double> MDSimulation::get_pressure_profile(int n_bins_local) {   // src/md_env.cu:3821-3918
    std::vector<double> out;
    if (n_bins_local <= 0) return out;
    out.assign(static_cast<size_t>(3 * n_bins_local), 0.0);
    check(mdx_pressure_profile(sim_, n_bins_local, out.data()), "get_pressure_profile");
    return out;
}

std::vector<double> MDSimulation::get_density_profile(int n_bins_per_rank) {
    const std::vector<int> n = get_N_profile(n_bins_per_rank);
    const size_t nb = n.size() / 2;
    std::vector<double> out(nb, 0.0);
    for (size_t i = 0; i < nb; ++i) {
        const double a = static_cast<double>(n[i]), b = static_cast<double>(n[i + nb]);
        const double sum = a + b;
        out[i] = sum == 0.0 ? 0.0 : (a - b) / sum;
    }
    return out;
}

void MDSimulation::sample_collect() {
    h_particles.resize(static_cast<size_t>(cfg_manager.config.n_particles_global));
    check(mdx_collect(sim_, reinterpret_cast<mdx_particle64 *>(h_particles.data()),
                      static_cast<int64_t>(h_particles.size())),
          "collect_particles_d2h");
}

void MDSimulation::get_host_particles(std::vector<Particle> &out) const { out = h_particles; }

void MDSimulation::save_env(const std::string &filename, const int step) {   // src/md_env.cu:4106-4124
    if (!particle_writer) {
        try {
            particle_writer = std::make_unique<FileWriter>(filename, true);
        } catch (const std::exception &) {
            particle_writer = std::make_unique<FileWriter>(filename, false);
        }
    }
    particle_writer->write_frame(h_particles.data(), h_particles.size(), static_cast<std::uint64_t>(step));
}

void MDSimulation::set_precision(int precision) {
    check(mdx_set_precision(sim_, precision), "set_precision", true);
    precision_ = precision;
}
