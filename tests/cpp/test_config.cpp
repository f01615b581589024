// Config round trip with the values of the reference's
// tests/md_config_load_save/test_config.cpp (one rank).  Writes argv[1],
// loads it back, and checks every field survives.
#include <cmath>
#include <cstdio>
#include <string>

#include "md_config.hpp"

int main(int argc, char **argv) {
    if (argc < 2) return 2;
    const int world_size = 1, world_rank = 0;
    MDConfig cfg{};
    cfg.n_particles_global = 2048;
    cfg.n_particles_type0 = 1024;
    cfg.box_w_global = std::sqrt(cfg.n_particles_global / 0.6 * 6);
    cfg.box_h_global = std::sqrt(cfg.n_particles_global / 0.6 / 6);
    cfg.T_init = 0.5;
    cfg.T_target = 1.0;
    cfg.EPSILON_AB = 0.25;
    cfg.dt = 1e-3;
    cfg.rank_size = world_size;
    cfg.rank_idx = 0;
    cfg.n_local = 0;
    cfg.n_halo_left = 0;
    cfg.n_halo_right = 0;
    cfg.n_cap = cfg.n_particles_global / world_size * 2;
    cfg.halo_left_cap = cfg.n_particles_global / world_size / 2;
    cfg.halo_right_cap = cfg.n_particles_global / world_size / 2;
    cfg.left_rank = (world_rank + world_size - 1) % world_size;
    cfg.right_rank = (world_rank + 1) % world_size;
    cfg.x_min = cfg.box_w_global / world_size * world_rank;
    cfg.x_max = cfg.box_w_global / world_size * (world_rank + 1);
    MDConfigManager mgr(cfg);
    mgr.config_to_json(argv[1]);
    MDConfigManager back = MDConfigManager::config_from_json(argv[1]);
    const MDConfig &b = back.config;
    int bad = 0;
    bad += b.box_w_global != cfg.box_w_global;
    bad += b.box_h_global != cfg.box_h_global;
    bad += b.x_max != cfg.x_max;
    bad += b.n_cap != cfg.n_cap || b.halo_left_cap != cfg.halo_left_cap;
    bad += b.run_name != "test" || b.load_name != "";
    bad += b.pressure_target_mode != "auto_initial";
    bad += back.serialize() != mgr.serialize();
    // overrides: --Dkey=value, --key=value, Dkey=value
    std::string k, v;
    bad += !MDConfigManager::parse_override_argument("--DT_target=0.75", k, v) || k != "T_target" || v != "0.75";
    bad += !back.apply_override(k, v) || back.config.T_target != 0.75;
    bad += !MDConfigManager::parse_override_argument("--n_particles_global=4096", k, v);
    bad += !back.apply_override(k, v) || back.config.n_particles_global != 4096;
    bad += back.apply_override("n_particles_global", "1e3");   // not an integer
    bad += back.apply_override("no_such_key", "1");
    bool threw = false;
    try {
        back.apply_overrides({{"dt", "abc"}});
    } catch (const std::exception &e) {
        threw = std::string(e.what()) == "Invalid override 'dt=abc'";
    }
    bad += !threw;
    if (argc > 2) back.print_config();
    std::printf("test_config bad=%d\n", bad);
    return bad == 0 ? 0 : 1;
}
