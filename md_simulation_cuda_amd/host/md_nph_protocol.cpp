// NphProtocol (phases and rules in md_nph_protocol.hpp).
#include "md_nph_protocol.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <numeric>

namespace {

struct MeanStd {
    double mean = 0.0, std = 0.0;
};

MeanStd mean_std(const std::deque<double> &v) {
    MeanStd m;
    if (v.empty()) return m;
    m.mean = std::accumulate(v.begin(), v.end(), 0.0) / static_cast<double>(v.size());
    if (v.size() > 1) {
        double s = 0.0;
        for (double x : v) s += (x - m.mean) * (x - m.mean);
        m.std = std::sqrt(s / static_cast<double>(v.size() - 1));
    }
    return m;
}

// Appends one formatted line to dir/name (rank 0 writes in the reference).
template <typename... A>
void append_line(const std::string &dir, const char *name, const char *fmt, A... a) {
    if (dir.empty()) return;
    const std::string path = (std::filesystem::path(dir) / name).string();
    if (std::FILE *f = std::fopen(path.c_str(), "a")) {
        std::fprintf(f, fmt, a...);
        std::fclose(f);
    }
}

}  // namespace

const char *NphProtocol::phase_name(Phase p) {
    switch (p) {
    case Phase::AUTO_TARGET: return "AUTO_TARGET";
    case Phase::NPH_WARMUP: return "NPH_WARMUP";
    case Phase::NPH_PRODUCTION: return "NPH_PRODUCTION";
    }
    return "UNKNOWN";
}

NphProtocol::NphProtocol(MDSimulation &sim, const MDConfig &cfg, const std::string &out_dir)
    : sim_(sim), cfg_(cfg), dir_(out_dir) {
    record_every_ = std::max(1, static_cast<int>(cfg.save_dt_interval / cfg.dt));
    eval_every_ = std::max(1, cfg.pressure_eval_interval_steps);
    auto_steps_ = std::max(1, cfg.pressure_target_auto_steps);
    window_ = std::max(2, cfg.pressure_window_samples);
    stable_needed_ = std::max(1, cfg.pressure_stable_required_windows);
    warm_min_ = std::max(0, cfg.nph_warmup_min_steps);
    warm_max_ = std::max<int64_t>(warm_min_ + 1, cfg.nph_warmup_max_steps);
    nph_target_ = std::max(1, cfg.nph_steps);
    snapshot_every_ = 20 * static_cast<int64_t>(record_every_);
    hard_limit_ = auto_steps_ + warm_max_ + nph_target_ + 50000;
    phase_ = cfg.pressure_target_mode == "auto_initial" ? Phase::AUTO_TARGET : Phase::NPH_WARMUP;
    P_target_ = cfg.P_target;
    switched_[static_cast<int>(phase_)] = 0;
    if (!dir_.empty()) {   // fresh logs with the reference's headers
        std::filesystem::create_directories(dir_);
        for (const char *name : {"pressure_log.csv", "barostat_state_log.csv", "thermo_log.csv"})
            std::fclose(std::fopen((std::filesystem::path(dir_) / name).string().c_str(), "w"));
        append_line(dir_, "pressure_log.csv", "%s",
                    "step,t,phase,P_inst,P_target,rolling_mean,rolling_std,rel_err,Lx,Ly,area,dA_dt\n");
        append_line(dir_, "barostat_state_log.csv", "%s",
                    "step,t,phase,Lx,Ly,area,dA_dt,a_barostat,P_inst,P_target\n");
        append_line(dir_, "thermo_log.csv", "%s", "step,t,phase,U_tot,K_tot,P_inst,P_target,Lx,Ly,area\n");
    }
}

bool NphProtocol::step() {
    if (finished_ || step_ >= hard_limit_) return false;
    const Phase used = phase_;
    double P_inst;
    if (used == Phase::AUTO_TARGET) {
        sim_.step_single_nose_hoover();
        P_inst = sim_.cal_instant_pressure();
    } else {
        sim_.step_single_NPH(P_target_);
        P_inst = sim_.get_last_instant_pressure();
    }
    ++done_[static_cast<int>(used)];
    const bool eval_now = step_ % eval_every_ == 0;
    if (eval_now) {
        if (used == Phase::AUTO_TARGET) {
            auto_samples_.push_back(P_inst);
        } else {
            window_samples_.push_back(P_inst);
            if (static_cast<int>(window_samples_.size()) > window_) window_samples_.pop_front();
        }
    }
    auto enter = [&](Phase p) {
        phase_ = p;
        switched_[static_cast<int>(p)] = step_;
    };
    if (used == Phase::AUTO_TARGET && done_[0] >= auto_steps_ && !auto_samples_.empty()) {
        P_target_ = mean_std(auto_samples_).mean;
        window_samples_.clear();
        stable_ = 0;
        enter(Phase::NPH_WARMUP);
    }
    if (used == Phase::NPH_WARMUP) {
        if (done_[1] >= warm_min_ && static_cast<int>(window_samples_.size()) >= window_) {
            const MeanStd w = mean_std(window_samples_);
            const double scale = std::max(1e-12, std::fabs(P_target_));
            const bool ok = std::fabs(w.mean - P_target_) / scale <= cfg_.pressure_tolerance_rel &&
                            w.std / scale <= cfg_.pressure_tolerance_rel;
            stable_ = ok ? stable_ + 1 : 0;
            if (stable_ >= stable_needed_) enter(Phase::NPH_PRODUCTION);
        }
        if (phase_ == Phase::NPH_WARMUP && done_[1] >= warm_max_) {
            forced_ = true;
            enter(Phase::NPH_PRODUCTION);
        }
    }
    log_rows(used, P_inst, eval_now, step_ % record_every_ == 0);
    if (used == Phase::NPH_PRODUCTION && done_[2] >= nph_target_) finished_ = true;
    ++step_;
    return !finished_ && step_ < hard_limit_;
}

int64_t NphProtocol::run() {
    const int64_t start = step_;
    while (step()) {
    }
    return step_ - start;
}

void NphProtocol::log_rows(Phase used, double P_inst, bool eval_now, bool record_now) {
    if (dir_.empty()) return;
    const double Lx = sim_.get_Lx(), Ly = sim_.get_Ly(), area = Lx * Ly;
    const bool nph = used != Phase::AUTO_TARGET;
    const double dA_dt = nph ? sim_.get_nph_area_rate() : 0.0;
    if (eval_now) {
        const MeanStd w = mean_std(window_samples_);
        const double rel_err = std::fabs(w.mean - P_target_) / std::max(1e-12, std::fabs(P_target_));
        const double a_baro = nph ? (P_inst - P_target_) / std::max(cfg_.barostat_mass, 1e-12) : 0.0;
        append_line(dir_, "pressure_log.csv", "%lld,%.6f,%s,%.10e,%.10e,%.10e,%.10e,%.10e,%.10e,%.10e,%.10e,%.10e\n",
                    static_cast<long long>(step_), sim_.t, phase_name(used), P_inst, P_target_, w.mean, w.std, rel_err,
                    Lx, Ly, area, dA_dt);
        append_line(dir_, "barostat_state_log.csv", "%lld,%.6f,%s,%.10e,%.10e,%.10e,%.10e,%.10e,%.10e,%.10e\n",
                    static_cast<long long>(step_), sim_.t, phase_name(used), Lx, Ly, area, dA_dt, a_baro, P_inst,
                    P_target_);
    }
    if (record_now) {
        sim_.sample_collect();
        const double U = sim_.cal_total_U(), K = sim_.cal_total_K();
        append_line(dir_, "thermo_log.csv", "%lld,%.6f,%s,%.10e,%.10e,%.10e,%.10e,%.10e,%.10e,%.10e\n",
                    static_cast<long long>(step_), sim_.t, phase_name(used), U, K, P_inst, P_target_, Lx, Ly, area);
        if (step_ % snapshot_every_ == 0)
            sim_.save_env((std::filesystem::path(dir_) / "saved_env.bin").string(), static_cast<int>(step_));
    }
}
