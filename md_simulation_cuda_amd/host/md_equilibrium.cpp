// EquilibriumTracker (declared, with its rules, in md_equilibrium.hpp).
#include "md_equilibrium.hpp"

#include <algorithm>
#include <cmath>

void EquilibriumTracker::reset(double interval, double dt) {
    interval_ = interval > 0.0 ? interval : (dt > 0.0 ? dt : 1.0);
    window_ = static_cast<std::size_t>(std::max(1L, std::lround(kWindowTime / interval_)));
    ring_.assign(2 * window_ * (kBasePasses + 1), 0.0);
    head_ = 0;
    count_ = 0;
    streak_ = 0;
}

double EquilibriumTracker::at(std::size_t i) const { return ring_[(head_ + i) % ring_.size()]; }

void EquilibriumTracker::record(double U) {
    if (ring_.empty()) reset(interval_, 0.0);
    if (count_ == ring_.size()) {   // drop the oldest
        ring_[head_] = U;
        head_ = (head_ + 1) % ring_.size();
    } else {
        ring_[(head_ + count_) % ring_.size()] = U;
        ++count_;
    }
}

bool EquilibriumTracker::evaluate(double sensitivity, int n_particles) {
    const std::size_t W = window_;
    if (W == 0 || count_ < 2 * W) return false;
    struct Win {
        double first, last, mean, var;
    };
    auto window_at = [&](std::size_t start) {
        Win w{at(start), at(start + W - 1), 0.0, 0.0};
        for (std::size_t i = 0; i < W; ++i) w.mean += at(start + i);
        w.mean /= static_cast<double>(W);
        if (W > 1) {
            for (std::size_t i = 0; i < W; ++i) {
                const double d = at(start + i) - w.mean;
                w.var += d * d;
            }
            w.var /= static_cast<double>(W - 1);
        }
        return w;
    };
    auto rel_change = [](const Win &w) {
        constexpr double tiny = 1e-12;
        double denom = w.first;
        if (std::fabs(denom) < tiny) denom = std::fabs(w.last) > tiny ? w.last : (w.first >= 0.0 ? 1.0 : -1.0);
        return std::fabs(denom) < tiny ? 0.0 : (w.last - w.first) / denom;
    };
    const Win a = window_at(count_ - 2 * W), b = window_at(count_ - W);
    const double tol = sensitivity / std::sqrt(static_cast<double>(std::max(1, n_particles)));
    const bool flat = std::fabs(rel_change(a)) < tol && std::fabs(rel_change(b)) < tol;
    const double se = std::sqrt(a.var / static_cast<double>(W) + b.var / static_cast<double>(W) + 1e-18);
    const double z = se > 0.0 ? std::fabs(a.mean - b.mean) / se : 0.0;
    const double alpha = std::clamp(kBaseAlpha / sensitivity, 1e-4, 0.5);
    const bool same_mean = std::erfc(z / std::sqrt(2.0)) > alpha;
    streak_ = (flat && same_mean) ? streak_ + 1 : 0;
    const int need = std::max(1, static_cast<int>(std::round(kBasePasses / sensitivity)));
    return streak_ >= need;
}
