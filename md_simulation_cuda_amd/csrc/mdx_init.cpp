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

// The constructor's starting frame (behaviour of src/md_env.cu:1296-1441, the
// rank-0 part), built from this specification:
//  - the box splits at x_split = devide_p * Lx into an A slab [0, x_split)
//    holding particles 0 .. n_A-1 (type 0) and a B slab [x_split, Lx) holding
//    n_A .. N-1 (type 1); afterwards every x moves by
//    Lx (1 - devide_p) / 2 and wraps into [0, Lx), which centres the A slab;
//  - each slab gets its own square lattice (Lattice below); particle k of a
//    slab sits at column k % cols, row k / cols;
//  - one std::mt19937 seeded with 12345 feeds every draw, in this order:
//    for each placed particle (A, then B) a U(-1e-5, 1e-5) jitter of x, then
//    of y; then for particle 0 .. N-1 vx, then vy from N(0, sqrt(T / m));
//    then theta ~ U(0, 2 pi) for particle 0 .. N-1.  A site outside its slab
//    is left unplaced and draws nothing (it cannot happen for a lattice from
//    Lattice, but the rule fixes the stream);
//  - the velocities lose their centre-of-mass velocity (total mass from the
//    two counts) and are scaled to K = (2N - 2) T / 2.
// The libstdc++ distributions define the stream, so they are used as such.

// Square lattice for `count` sites in a w x h slab: the pitch starts at
// sqrt(w h / count) and is multiplied by 0.99 until floor(w / pitch) columns
// times floor(h / pitch) rows (each at least 1) hold `count` sites.
struct Lattice {
    double pitch;
    int64_t cols, rows;
};

Lattice fit_lattice(double w, double h, int64_t count) {
    Lattice L{std::sqrt(w * h / static_cast<double>(count)), 1, 1};
    for (;;) {
        L.cols = std::max<int64_t>(1, static_cast<int64_t>(w / L.pitch));
        L.rows = std::max<int64_t>(1, static_cast<int64_t>(h / L.pitch));
        if (L.cols * L.rows >= count) return L;
        L.pitch *= 0.99;
    }
}

struct Slab {
    double x0, x1;    // x range before the centring shift
    int64_t first;    // index of the slab's first particle
    int64_t count;
    int32_t type;
};

void place_slab(const Slab &s, double Lx, double Ly, double shift, std::mt19937 &rng, mdx_particle64 *P) {
    if (s.count <= 0) return;
    const Lattice L = fit_lattice(s.x1 - s.x0, Ly, s.count);
    std::uniform_real_distribution<double> jitter(-1.0e-5, 1.0e-5);
    for (int64_t k = 0; k < s.count; ++k) {
        const double col = static_cast<double>(k % L.cols), row = static_cast<double>(k / L.cols);
        const double x = (col + 0.5) * L.pitch + s.x0;
        const double y = (row + 0.5) * L.pitch;
        if (x >= s.x1 || y >= Ly) continue;
        mdx_particle64 &p = P[s.first + k];
        p.pos[0] = std::fmod(x + shift + Lx, Lx) + jitter(rng);
        p.pos[1] = y + jitter(rng);
        p.type = s.type;
    }
}

void init_particles(const mdx_init_params &ip, mdx_particle64 *P) {
    const int64_t N = ip.n_particles_global, nA = ip.n_particles_type0, nB = N - nA;
    const double Lx = ip.box_w, Ly = ip.box_h, T = ip.T_init;
    const double x_split = Lx * ip.devide_p;
    const double shift = Lx * (0.5 - ip.devide_p / 2.0);
    std::mt19937 rng(12345);
    // (the slab starting at 0 adds x0 = 0, which leaves every x unchanged)
    place_slab({0.0, x_split, 0, nA, 0}, Lx, Ly, shift, rng, P);
    place_slab({x_split, Lx, nA, nB, 1}, Lx, Ly, shift, rng, P);

    auto mass = [&](const mdx_particle64 &p) { return p.type == 0 ? ip.mass_a : ip.mass_b; };
    // Two draws per particle use up exactly the pair that each polar-method
    // step of normal_distribution makes, so one standard normal scaled by
    // sqrt(T / m) gives the same values as a fresh N(0, sqrt(T / m)) per
    // particle.
    std::normal_distribution<double> z;
    double px = 0.0, py = 0.0;
    for (int64_t i = 0; i < N; ++i) {
        const double m = mass(P[i]), s = std::sqrt(T / m);
        P[i].vel[0] = z(rng) * s;
        P[i].vel[1] = z(rng) * s;
        px += m * P[i].vel[0];
        py += m * P[i].vel[1];
    }
    const double M = static_cast<double>(nA) * ip.mass_a + static_cast<double>(nB) * ip.mass_b;
    const double ux = px / M, uy = py / M;
    double K = 0.0;
    for (int64_t i = 0; i < N; ++i) {
        P[i].vel[0] -= ux;
        P[i].vel[1] -= uy;
        K += 0.5 * mass(P[i]) * (P[i].vel[0] * P[i].vel[0] + P[i].vel[1] * P[i].vel[1]);
    }
    const double scale = std::sqrt((2.0 * static_cast<double>(N) - 2.0) * 0.5 * T / K);
    std::uniform_real_distribution<double> angle(0.0, 2.0 * M_PI);
    for (int64_t i = 0; i < N; ++i) {
        P[i].vel[0] *= scale;
        P[i].vel[1] *= scale;
    }
    for (int64_t i = 0; i < N; ++i) P[i].theta = angle(rng);
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
