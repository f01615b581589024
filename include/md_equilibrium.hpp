// Equilibrium test on a record of the potential energy: the statistics behind
// MDSimulation::check_eqlibrium (behaviour of src/md_env.cu:174-313 and
// 2478-2500 of the reference), as a class of its own so that it can be driven
// by any series.
//
// U is recorded every `interval` time units; a window holds the samples of
// 100 time units (round(100 / interval), at least 1).  A test needs two full
// windows, the latest W samples and the W before them, and passes when
//   - in each window the relative change (U_last - U_first) / U_first (with
//     U_last, or the sign of U_first, as denominator when |U_first| < 1e-12)
//     is below sensitivity / sqrt(N) in magnitude, and
//   - the two windows' means agree: with their sample variances v1, v2,
//     z = |m1 - m2| / sqrt(v1 / W + v2 / W + 1e-18) has the two-sided normal
//     tail erfc(z / sqrt 2) above alpha = clamp(0.2 / sensitivity, 1e-4, 0.5).
// A passing test extends a streak, a failing one resets it, a test without
// two full windows leaves it alone; the system counts as equilibrated while
// the streak is at least max(1, round(3 / sensitivity)).  The record keeps the
// latest 8 W samples.
#pragma once

#include <cstddef>
#include <vector>

class EquilibriumTracker {
public:
    static constexpr double kWindowTime = 100.0;
    static constexpr int kBasePasses = 3;
    static constexpr double kBaseAlpha = 0.2;

    // record interval: `interval` if positive, else dt if positive, else 1
    void reset(double interval, double dt);
    double interval() const { return interval_; }
    std::size_t window() const { return window_; }
    std::size_t size() const { return count_; }
    int streak() const { return streak_; }

    void record(double U);
    // One test at this sensitivity (the caller clamps it away from 0) for a
    // system of n_particles; true when the streak reaches the required count.
    bool evaluate(double sensitivity, int n_particles);

private:
    double at(std::size_t i) const;   // i-th oldest sample kept
    double interval_ = 1.0;
    std::size_t window_ = 1;
    std::vector<double> ring_;
    std::size_t head_ = 0, count_ = 0;
    int streak_ = 0;
};
