// MDConfigManager: JSON load/store, serialize, overrides.
// Restates src/md_config.cpp:43-495 of the reference with json_lite in place
// of nlohmann::json.
#include "md_config.hpp"

#include <algorithm>
#include <cctype>
#include <charconv>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <type_traits>

#include "json_lite.hpp"

using mdx_json::Value;

namespace {

template <typename T>
bool parse_scalar(const std::string &text, T &out) {   // src/md_config.cpp:9-18
    std::istringstream iss(text);
    iss >> out;
    if (iss.fail()) return false;
    iss >> std::ws;
    return iss.eof();
}

template <>
bool parse_scalar<std::string>(const std::string &text, std::string &out) {
    out = text;
    return true;
}

// j.value(key, default) of nlohmann: the stored value converted to the
// default's type, or the default when the key is absent.
template <typename T>
T value_or(const Value &j, const char *key, const T &def) {
    if (!j.contains(key)) return def;
    const Value &v = j.at(key);
    if constexpr (std::is_same_v<T, std::string>) {
        return v.as_string();
    } else if constexpr (std::is_integral_v<T>) {
        return static_cast<T>(v.as_int());
    } else {
        return static_cast<T>(v.as_double());
    }
}

// One table drives load, store, override and print so the key set cannot
// drift between them (src/md_config.cpp:43-199, 329-458).
template <typename F>
void for_each_field(MDConfig &c, F &&f) {
    f("n_particles_global", c.n_particles_global);
    f("n_particles_type0", c.n_particles_type0);
    f("box_w_global", c.box_w_global);
    f("box_h_global", c.box_h_global);
    f("T_init", c.T_init);
    f("T_target", c.T_target);
    f("SIGMA_AA", c.SIGMA_AA);
    f("SIGMA_BB", c.SIGMA_BB);
    f("SIGMA_AB", c.SIGMA_AB);
    f("EPSILON_AA", c.EPSILON_AA);
    f("EPSILON_BB", c.EPSILON_BB);
    f("EPSILON_AB", c.EPSILON_AB);
    f("MASS_A", c.MASS_A);
    f("MASS_B", c.MASS_B);
    f("devide_p", c.devide_p);
    f("dt", c.dt);
    f("Q", c.Q);
    f("save_dt_interval", c.save_dt_interval);
    f("cutoff", c.cutoff);
    f("run_name", c.run_name);
    f("load_name", c.load_name);
    f("THREADS_PER_BLOCK", c.THREADS_PER_BLOCK);
    f("rank_size", c.rank_size);
    f("mu", c.mu);
    f("v0", c.v0);
    f("D_r", c.D_r);
    f("D_theta", c.D_theta);
    f("barostat_mass", c.barostat_mass);
    f("k_piston", c.k_piston);
    f("barostat_area_rate_init", c.barostat_area_rate_init);
    f("barostat_height_min_ratio", c.barostat_height_min_ratio);
    f("barostat_height_max_ratio", c.barostat_height_max_ratio);
    f("pressure_target_mode", c.pressure_target_mode);
    f("P_target", c.P_target);
    f("pressure_target_auto_steps", c.pressure_target_auto_steps);
    f("pressure_eval_interval_steps", c.pressure_eval_interval_steps);
    f("pressure_window_samples", c.pressure_window_samples);
    f("pressure_tolerance_rel", c.pressure_tolerance_rel);
    f("pressure_stable_required_windows", c.pressure_stable_required_windows);
    f("nph_warmup_min_steps", c.nph_warmup_min_steps);
    f("nph_warmup_max_steps", c.nph_warmup_max_steps);
    f("nph_steps", c.nph_steps);
    f("rank_idx", c.rank_idx);
    f("n_local", c.n_local);
    f("n_halo_left", c.n_halo_left);
    f("n_halo_right", c.n_halo_right);
    f("n_cap", c.n_cap);
    f("halo_left_cap", c.halo_left_cap);
    f("halo_right_cap", c.halo_right_cap);
    f("left_rank", c.left_rank);
    f("right_rank", c.right_rank);
    f("x_min", c.x_min);
    f("x_max", c.x_max);
}

void load_from_json_object(const Value &j, MDConfig &cfg) {
    // Start from the struct defaults (value-initialised), as the reference
    // does; each field is its own default when the key is absent.
    MDConfig out{};
    for_each_field(out, [&](const char *key, auto &field) {
        using T = std::decay_t<decltype(field)>;
        field = value_or<T>(j, key, field);
    });
    cfg = out;
}

void store_to_json_object(Value &j, MDConfig cfg) {
    for_each_field(cfg, [&](const char *key, auto &field) {
        using T = std::decay_t<decltype(field)>;
        if constexpr (std::is_same_v<T, std::string>)
            j[key] = Value::from_string(field);
        else if constexpr (std::is_integral_v<T>)
            j[key] = Value::from_int(field);
        else
            j[key] = Value::from_double(field);
    });
}

// {} formatting of fmt for the printed config: shortest round-trip doubles.
std::string fmt_value(double v) {
    char buf[64];
    auto r = std::to_chars(buf, buf + sizeof(buf), v);
    return std::string(buf, r.ptr);
}
std::string fmt_value(int v) { return std::to_string(v); }
std::string fmt_value(const std::string &v) { return v; }

}  // namespace

MDConfigManager::MDConfigManager(MDConfig config) { this->config = config; }

MDConfigManager::~MDConfigManager() = default;

void MDConfigManager::print_config() {
    std::string out = "MDConfig:\n";
    MDConfig c = config;
    for_each_field(c, [&](const char *key, auto &field) {
        out += "  ";
        out += key;
        out += ": ";
        out += fmt_value(field);
        out += "\n";
    });
    std::fputs(out.c_str(), stdout);
    std::fflush(stdout);
}

MDConfigManager MDConfigManager::config_from_json(const std::string &filepath) {
    std::ifstream in(filepath);
    if (!in.is_open()) throw std::runtime_error("Failed to open config file: " + filepath);
    std::stringstream ss;
    ss << in.rdbuf();
    const Value j = mdx_json::parse(ss.str());
    MDConfig cfg{};
    load_from_json_object(j, cfg);
    return MDConfigManager(cfg);
}

void MDConfigManager::config_to_json(const std::string &filepath) {
    Value j = Value::object();
    store_to_json_object(j, config);
    std::ofstream out(filepath);
    if (!out.is_open()) throw std::runtime_error("Failed to open config file for writing: " + filepath);
    out << j.dump(4) << std::endl;
}

std::string MDConfigManager::serialize() const {
    Value j = Value::object();
    store_to_json_object(j, config);
    return j.dump();
}

void MDConfigManager::deserialize(const std::string &data) {
    const Value j = mdx_json::parse(data);
    load_from_json_object(j, config);
}

bool MDConfigManager::apply_override(const std::string &key, const std::string &value) {
    bool found = false, ok = false;
    for_each_field(config, [&](const char *k, auto &field) {
        if (found || key != k) return;
        found = true;
        using T = std::decay_t<decltype(field)>;
        T parsed{};
        if (parse_scalar<T>(value, parsed)) {
            field = parsed;
            ok = true;
        }
    });
    return found && ok;
}

void MDConfigManager::apply_overrides(const std::vector<MDConfigOverride> &overrides) {
    for (const auto &o : overrides)
        if (!apply_override(o.key, o.value))
            throw std::runtime_error("Invalid override '" + o.key + "=" + o.value + "'");
}

// src/md_config.cpp:470-495: strips a leading "--", then a leading 'D', then
// splits at the first '='.
bool MDConfigManager::parse_override_argument(const std::string &argument, std::string &out_key,
                                              std::string &out_value) {
    if (argument.empty()) return false;
    std::string token = argument;
    if (token.rfind("--", 0) == 0) token.erase(0, 2);
    if (token.rfind('D', 0) == 0) token.erase(0, 1);
    const auto eq = token.find('=');
    if (eq == std::string::npos) return false;
    out_key = token.substr(0, eq);
    out_value = token.substr(eq + 1);
    return !out_key.empty();
}
