// MDConfig / MDConfigManager of the reference (include/md_config.hpp:14-111
// there), without the nlohmann and fmt dependencies.  Keys, defaults, the
// JSON layout of config_to_json, serialize/deserialize and the override
// parser follow src/md_config.cpp:43-495.
#pragma once

#include <string>
#include <vector>

struct MDConfig {
    // Global params for all ranks
    int n_particles_global;
    int n_particles_type0;
    double box_w_global;
    double box_h_global;
    double T_init;
    double T_target;
    double SIGMA_AA = 1.0;
    double SIGMA_BB = 1.0;
    double SIGMA_AB = 1.0;
    double EPSILON_AA = 1.0;
    double EPSILON_BB = 1.0;
    double EPSILON_AB = 0.25;
    double MASS_A = 1.0;
    double MASS_B = 1.0;
    double devide_p = 0.5;
    double dt = 1e-3;
    double Q = 100.0;
    double save_dt_interval = 0.1;
    double cutoff = 2.5;
    std::string run_name = "test";
    std::string load_name = "";
    int THREADS_PER_BLOCK = 256;
    int rank_size = 8;

    // ABP parameters (round-tripped; ABP dynamics are not part of this library)
    double mu = 1.0;
    double v0 = 0.0;
    double D_r = 0.0;
    double D_theta = 0.0;

    // Periodic-box NPH barostat controls
    double barostat_mass = 1.0e4;
    double k_piston = 100.0;
    double barostat_area_rate_init = 0.0;
    double barostat_height_min_ratio = 0.6;
    double barostat_height_max_ratio = 1.4;
    std::string pressure_target_mode = "auto_initial";
    double P_target = 0.0;
    int pressure_target_auto_steps = 5000;
    int pressure_eval_interval_steps = 100;
    int pressure_window_samples = 50;
    double pressure_tolerance_rel = 0.05;
    int pressure_stable_required_windows = 3;
    int nph_warmup_min_steps = 20000;
    int nph_warmup_max_steps = 200000;
    int nph_steps = 300000;

    // Per-rank fields: accepted and round-tripped; one GPU holds the whole box.
    int rank_idx = 0;
    int n_local;
    int n_halo_left;
    int n_halo_right;
    int n_cap;
    int halo_left_cap;
    int halo_right_cap;
    int left_rank;
    int right_rank;
    double x_min;
    double x_max;
};

struct MDConfigOverride {
    std::string key;
    std::string value;
};

class MDConfigManager {
public:
    MDConfigManager() = default;
    MDConfigManager(MDConfig config);
    ~MDConfigManager();

    struct MDConfig config;

    static MDConfigManager config_from_json(const std::string &filepath);
    void config_to_json(const std::string &filepath);
    std::string serialize() const;
    void deserialize(const std::string &data);
    void print_config();
    bool apply_override(const std::string &key, const std::string &value);
    void apply_overrides(const std::vector<MDConfigOverride> &overrides);
    static bool parse_override_argument(const std::string &argument, std::string &out_key,
                                        std::string &out_value);
};
