// MDSimulation: the environment API of the reference (include/md_env.hpp:51-263
// there) for the 2D binary-LJ step, on one GPU through the mdx C ABI.
//
// The MPI communicator of the reference's constructors is accepted and
// ignored (one GPU holds the whole periodic box).  Triangulation, interface,
// profile and ABP members of the reference are not part of this library.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "md_common.hpp"
#include "md_config.hpp"
#include "md_equilibrium.hpp"
#include "md_particle.hpp"
#include "mdx.h"

class MDSimulation {
public:
    explicit MDSimulation(MDConfigManager config_manager);
    MDSimulation(MDConfigManager config_manager, const std::string &filename, int step);
    // Reference signatures with a communicator (src/md_env.cu:21, :81).
    template <class Comm>
    MDSimulation(MDConfigManager config_manager, Comm) : MDSimulation(std::move(config_manager)) {}
    template <class Comm>
    MDSimulation(MDConfigManager config_manager, Comm, const std::string &filename, int step)
        : MDSimulation(std::move(config_manager), filename, step) {}
    ~MDSimulation();
    MDSimulation(const MDSimulation &) = delete;
    MDSimulation &operator=(const MDSimulation &) = delete;

    double t = 0.0;   // evolution time

    void save_env(const std::string &filename, const int step);
    void get_host_particles(std::vector<Particle> &out) const;

    double cal_total_K();
    double cal_total_U();
    double cal_partial_U_lambda(double epsilon_lambda);
    // counts of type 0, then of other types, in n_bins_per_rank slabs along x
    // (src/md_env.cu:3920-3988 with one rank: 2 * n_bins_per_rank values)
    std::vector<int> get_N_profile(int n_bins_per_rank);
    // per slab (n_A - n_B) / (n_A + n_B), 0 for an empty slab (src/md_env.cu:4086-4104)
    std::vector<double> get_density_profile(int n_bins_per_rank);
    // Pxx, Pyy, Pxy in n_bins_local slabs along y, each divided by the slab's
    // area (src/md_env.cu:3821-3918 with one rank: 3 * n_bins_local values;
    // empty for n_bins_local <= 0)
    std::vector<double> get_pressure_profile(int n_bins_local);
    double get_Lx() const { return cfg_manager.config.box_w_global; }
    double get_Ly() const { return cfg_manager.config.box_h_global; }

    void middle_reflect_LG();   // src/md_env.cu:1457-1508
    void step_single();   // README name of step_single_NVE
    void step_single_NVE();
    void step_single_nose_hoover(bool do_middle_wrap = false);
    void step_single_NPH(double pressure_ext);
    void step_single_NPH_piston(double pressure_ext);
    double cal_instant_pressure();
    double get_last_instant_pressure() const { return last_instant_pressure; }
    double get_nph_area_rate() const;
    double get_nph_piston_velocity() const { return nph_piston_velocity; }
    double get_nph_piston_force() const { return nph_piston_force; }

    void sample_collect();
    // Records U at every save_dt_interval boundary t has passed since the last
    // call (cal_total_U at the current state), then runs one equilibrium test
    // (EquilibriumTracker) at max(sensitivity, 1e-3).  The reference's name
    // and spelling (src/md_env.cu:2478-2500).
    bool check_eqlibrium(double sensitivity);

    // Extensions of this library.
    void step_many(int kind, int nsteps, double pressure_ext = 0.0, bool do_middle_wrap = false);
    void set_precision(int precision);   // mdx_precision
    // Tuning of the pair pass (mdx_tuning).  By default (environment variable
    // MDX_TUNING unset or "auto") the first NVE or Nose-Hoover step runs the
    // autotuner on a copy of the system and keeps its choice; MDX_TUNING="off"
    // keeps the cell list every step; "cell_scale,lanes,skin,window" sets a
    // tuning.  Results differ between tunings only by FP32 summation order.
    void set_tuning(double cell_scale, int lanes_per_particle, double skin, int window);
    void autotune(int kind, int steps_per_candidate = 20);
    mdx_sim *handle() { return sim_; }
    const MDConfigManager &config() const { return cfg_manager; }

private:
    void create_from_frame(const std::vector<Particle> &frame);
    void sync_scalars();
    void check(int rc, const char *where, bool throws = false) const;

    MDConfigManager cfg_manager;
    mdx_sim *sim_ = nullptr;
    std::unique_ptr<FileWriter> particle_writer;
    std::vector<Particle> h_particles;
    double last_instant_pressure = 0.0;
    double nph_area_momentum = 0.0;
    double nph_piston_velocity = 0.0;
    double nph_piston_force = 0.0;
    int precision_ = MDX_PREC_COMP32;
    bool autotune_pending_ = false;   // MDX_TUNING auto: tune at the first NVE / NVT step
    EquilibriumTracker eq_;           // U record of check_eqlibrium (reset by the constructors)
    double next_record_time_ = 0.0;
};
