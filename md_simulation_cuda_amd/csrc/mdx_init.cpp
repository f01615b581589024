// Initial frames: the reference constructor's init_particles and the seeded
// synthetic generators of the benchmark workloads.  Host code only.
#include "mdx.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

namespace {

// Restates MDSimulation::init_particles, src/md_env.cu:1296-1441 (rank 0 part).
void init_particles(const mdx_init_params &ip, mdx_particle64 *P) {
    const double devide_p = ip.devide_p;
    const int n_particles_global = static_cast<int>(ip.n_particles_global);
    const int n_particles_type0 = static_cast<int>(ip.n_particles_type0);
    const int n_particles_type1 = n_particles_global - n_particles_type0;
    const double box_w = ip.box_w, box_h = ip.box_h;
    const double unshifted_x_devide = box_w * devide_p;
    const double shift_dx = box_w * (0.5 - devide_p / 2.0);
    const double T_init = ip.T_init;
    const double mass_type0 = ip.mass_a, mass_type1 = ip.mass_b;
    std::mt19937 rng(12345);
    std::uniform_real_distribution<double> pos_perturb(-1.0e-5, 1.0e-5);
    double spacing_type0 = std::sqrt(unshifted_x_devide * box_h / n_particles_type0);
    double spacing_type1 = std::sqrt((box_w - unshifted_x_devide) * box_h / n_particles_type1);
    auto adjust_grid = [](double width, double height, int n, double &spacing, int &n_cols, int &n_rows) {
        n_rows = static_cast<int>(height / spacing);
        n_cols = static_cast<int>(width / spacing);
        if (n_rows <= 0) n_rows = 1;
        if (n_cols <= 0) n_cols = 1;
        while (n_rows * n_cols < n) {
            spacing *= 0.99;
            n_rows = static_cast<int>(height / spacing);
            n_cols = static_cast<int>(width / spacing);
            if (n_rows <= 0) n_rows = 1;
            if (n_cols <= 0) n_cols = 1;
        }
    };
    int n_rows_type0 = 0, n_cols_type0 = 0, n_rows_type1 = 0, n_cols_type1 = 0;
    adjust_grid(unshifted_x_devide, box_h, n_particles_type0, spacing_type0, n_cols_type0, n_rows_type0);
    adjust_grid(box_w - unshifted_x_devide, box_h, n_particles_type1, spacing_type1, n_cols_type1, n_rows_type1);
    for (int j = 0; j < n_rows_type0; ++j) {
        for (int i = 0; i < n_cols_type0; ++i) {
            const int idx = j * n_cols_type0 + i;
            if (idx >= n_particles_type0) break;
            const double x_local = (i + 0.5) * spacing_type0;
            const double y = (j + 0.5) * spacing_type0;
            if (x_local >= unshifted_x_devide) continue;
            if (y >= box_h) continue;
            const double x = std::fmod(x_local + shift_dx + box_w, box_w);
            P[idx].pos[0] = x + pos_perturb(rng);
            P[idx].pos[1] = y + pos_perturb(rng);
            P[idx].type = 0;
        }
    }
    for (int j = 0; j < n_rows_type1; ++j) {
        for (int i = 0; i < n_cols_type1; ++i) {
            const int idx = j * n_cols_type1 + i + n_particles_type0;
            if (idx >= n_particles_global) break;
            const double x_local = (i + 0.5) * spacing_type1 + unshifted_x_devide;
            const double y = (j + 0.5) * spacing_type1;
            if (x_local >= box_w) continue;
            if (y >= box_h) continue;
            const double x = std::fmod(x_local + shift_dx + box_w, box_w);
            P[idx].pos[0] = x + pos_perturb(rng);
            P[idx].pos[1] = y + pos_perturb(rng);
            P[idx].type = 1;
        }
    }
    const double total_mass = n_particles_type0 * mass_type0 + n_particles_type1 * mass_type1;
    double sum_px = 0.0, sum_py = 0.0;
    for (int i = 0; i < n_particles_global; ++i) {
        const double mass = (P[i].type == 0) ? mass_type0 : mass_type1;
        const double stddev = std::sqrt(T_init / mass);
        std::normal_distribution<double> dist(0.0, stddev);
        P[i].vel[0] = dist(rng);
        P[i].vel[1] = dist(rng);
        sum_px += mass * P[i].vel[0];
        sum_py += mass * P[i].vel[1];
    }
    const double vcm_x = sum_px / total_mass, vcm_y = sum_py / total_mass;
    for (int i = 0; i < n_particles_global; ++i) {
        P[i].vel[0] -= vcm_x;
        P[i].vel[1] -= vcm_y;
    }
    double current_K = 0.0;
    const double target_K = (2 * static_cast<double>(n_particles_global) - 2) * 0.5 * T_init;
    for (int i = 0; i < n_particles_global; ++i) {
        const double mass = (P[i].type == 0) ? mass_type0 : mass_type1;
        current_K += 0.5 * mass * (P[i].vel[0] * P[i].vel[0] + P[i].vel[1] * P[i].vel[1]);
    }
    const double rescale = std::sqrt(target_K / current_K);
    for (int i = 0; i < n_particles_global; ++i) {
        P[i].vel[0] *= rescale;
        P[i].vel[1] *= rescale;
    }
    std::uniform_real_distribution<double> theta_dist(0.0, 2.0 * M_PI);
    for (int i = 0; i < n_particles_global; ++i) P[i].theta = theta_dist(rng);
}

// Portable seeded streams: 53-bit uniforms from mt19937_64, Box-Muller normals.
struct Rng {
    std::mt19937_64 g;
    explicit Rng(uint64_t seed) : g(seed) {}
    double uniform() { return static_cast<double>(g() >> 11) * (1.0 / 9007199254740992.0); }
    double normal() {
        double u1 = uniform();
        while (u1 <= 0.0) u1 = uniform();
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

void assign_types(Rng &r, int64_t n, int64_t n0, mdx_particle64 *P) {
    std::vector<int32_t> t(n);
    for (int64_t i = 0; i < n; ++i) t[i] = i < n0 ? 0 : 1;
    for (int64_t i = n - 1; i > 0; --i) {   // Fisher-Yates
        const int64_t j = static_cast<int64_t>(r.uniform() * static_cast<double>(i + 1));
        std::swap(t[i], t[std::min(j, i)]);
    }
    for (int64_t i = 0; i < n; ++i) P[i].type = t[i];
}

void maxwell(Rng &r, int64_t n, double T, double ma, double mb, mdx_particle64 *P) {
    for (int64_t i = 0; i < n; ++i) P[i].vel[0] = P[i].vel[1] = 0.0;
    if (!(T > 0.0) || n < 2) return;
    double M = 0.0, px = 0.0, py = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        const double m = P[i].type == 0 ? ma : mb;
        const double s = std::sqrt(T / m);
        P[i].vel[0] = s * r.normal();
        P[i].vel[1] = s * r.normal();
        M += m;
        px += m * P[i].vel[0];
        py += m * P[i].vel[1];
    }
    double K = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        const double m = P[i].type == 0 ? ma : mb;
        P[i].vel[0] -= px / M;
        P[i].vel[1] -= py / M;
        K += 0.5 * m * (P[i].vel[0] * P[i].vel[0] + P[i].vel[1] * P[i].vel[1]);
    }
    const double f = std::sqrt((2.0 * static_cast<double>(n) - 2.0) * 0.5 * T / K);
    for (int64_t i = 0; i < n; ++i) {
        P[i].vel[0] *= f;
        P[i].vel[1] *= f;
    }
}

}  // namespace

extern "C" int mdx_init_particles(const mdx_init_params *ip, mdx_particle64 *out) {
    if (!ip || !out || ip->n_particles_global <= 0 || ip->n_particles_type0 < 0 ||
        ip->n_particles_type0 > ip->n_particles_global || !(ip->box_w > 0.0) || !(ip->box_h > 0.0))
        return MDX_ERR_INVALID;
    std::memset(out, 0, sizeof(mdx_particle64) * ip->n_particles_global);
    init_particles(*ip, out);
    return MDX_OK;
}

extern "C" int mdx_gen_lattice(int64_t n, int64_t n_type0, double density, double jitter, double T,
                               double mass_a, double mass_b, uint64_t seed, mdx_particle64 *out,
                               double *box_out) {
    if (!out || n <= 0 || n_type0 < 0 || n_type0 > n || !(density > 0.0)) return MDX_ERR_INVALID;
    std::memset(out, 0, sizeof(mdx_particle64) * n);
    const double L = std::sqrt(static_cast<double>(n) / density);
    int64_t side = static_cast<int64_t>(std::ceil(std::sqrt(static_cast<double>(n))));
    while (side * side < n) ++side;
    const double a = L / static_cast<double>(side);
    Rng r(seed);
    for (int64_t k = 0; k < n; ++k) {
        const int64_t i = k % side, j = k / side;
        double x = (static_cast<double>(i) + 0.5) * a + jitter * (2.0 * r.uniform() - 1.0);
        double y = (static_cast<double>(j) + 0.5) * a + jitter * (2.0 * r.uniform() - 1.0);
        x = std::fmod(x + L, L);
        y = std::fmod(y + L, L);
        out[k].pos[0] = x;
        out[k].pos[1] = y;
    }
    assign_types(r, n, n_type0, out);
    maxwell(r, n, T, mass_a, mass_b, out);
    if (box_out) {
        box_out[0] = L;
        box_out[1] = L;
    }
    return MDX_OK;
}

extern "C" int mdx_gen_uniform(int64_t n, int64_t n_type0, double density, double T, double mass_a,
                               double mass_b, uint64_t seed, mdx_particle64 *out, double *box_out) {
    if (!out || n <= 0 || n_type0 < 0 || n_type0 > n || !(density > 0.0)) return MDX_ERR_INVALID;
    std::memset(out, 0, sizeof(mdx_particle64) * n);
    const double L = std::sqrt(static_cast<double>(n) / density);
    Rng r(seed);
    for (int64_t k = 0; k < n; ++k) {
        out[k].pos[0] = r.uniform() * L;
        out[k].pos[1] = r.uniform() * L;
    }
    assign_types(r, n, n_type0, out);
    maxwell(r, n, T, mass_a, mass_b, out);
    if (box_out) {
        box_out[0] = L;
        box_out[1] = L;
    }
    return MDX_OK;
}
