// The three-phase NPH protocol of the reference's NPH test driver
// (tests/run_NPH_test/run_test_NPH.cpp:15-200 there; docs/20260303_NPH.md
// sections 340-505) as a reusable driver over MDSimulation.
//
//   AUTO_TARGET     Nose-Hoover steps; every pressure_eval_interval_steps the
//                   instantaneous pressure is sampled; after
//                   pressure_target_auto_steps steps P_target becomes the
//                   samples' mean (skipped unless pressure_target_mode is
//                   "auto_initial": then P_target is the config's).
//   NPH_WARMUP      NPH steps at P_target; the last pressure_window_samples
//                   samples (one every pressure_eval_interval_steps) form a
//                   window; once nph_warmup_min_steps have run and the window
//                   is full, each evaluation step checks |mean - P_target| and
//                   the sample standard deviation against
//                   pressure_tolerance_rel * max(|P_target|, 1e-12);
//                   pressure_stable_required_windows passes in a row, or
//                   nph_warmup_max_steps steps, end the phase.
//   NPH_PRODUCTION  nph_steps NPH steps.
// The run stops after the production steps, or at auto + warmup_max +
// nph_steps + 50000 steps.  Counts are clamped as the reference clamps them
// (window >= 2, intervals >= 1, warmup_max > warmup_min).
//
// With an output directory the driver writes the reference's three logs
// there (pressure_log.csv, barostat_state_log.csv and thermo_log.csv, same
// columns), thermo rows every save_dt_interval / dt steps, pressure and
// barostat rows every evaluation step, and, every 20 records, the frame to
// saved_env.bin through save_env.
#pragma once

#include <cstdint>
#include <deque>
#include <string>

#include "md_env.hpp"

class NphProtocol {
public:
    enum class Phase { AUTO_TARGET = 0, NPH_WARMUP = 1, NPH_PRODUCTION = 2 };
    static const char *phase_name(Phase p);

    // out_dir empty: no files
    NphProtocol(MDSimulation &sim, const MDConfig &cfg, const std::string &out_dir = "");

    // One step of the protocol; false once it has finished.
    bool step();
    // Steps until finished (or the hard limit); returns the steps taken.
    int64_t run();

    Phase phase() const { return phase_; }
    bool finished() const { return finished_; }
    double P_target() const { return P_target_; }
    int64_t steps_done() const { return step_; }
    int64_t steps_in(Phase p) const { return done_[static_cast<int>(p)]; }
    int64_t switch_step(Phase p) const { return switched_[static_cast<int>(p)]; }   // -1 if never entered
    bool warmup_forced() const { return forced_; }

private:
    void log_rows(Phase used, double P_inst, bool eval_now, bool record_now);

    MDSimulation &sim_;
    MDConfig cfg_;
    std::string dir_;
    int record_every_, eval_every_, auto_steps_, window_, stable_needed_;
    int64_t warm_min_, warm_max_, nph_target_, snapshot_every_, hard_limit_;
    Phase phase_;
    double P_target_;
    int64_t step_ = 0;
    int64_t done_[3] = {0, 0, 0};
    int64_t switched_[3] = {-1, -1, -1};
    int stable_ = 0;
    bool finished_ = false, forced_ = false;
    std::deque<double> auto_samples_, window_samples_;
};
