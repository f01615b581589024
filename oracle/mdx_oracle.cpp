// FP64 CPU oracle for the 2D binary-LJ MD step -- TEST INFRASTRUCTURE ONLY.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// this library, as the checker (or the timed CPU baseline); the shipped
// library never links it.  Every function restates the reference file:line it
// cites (paths relative to the reference repository).
#include "mdx_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct Sim {
    orc_params p;
    std::vector<orc_particle> P;
    double xi = 0.0;          // src/md_env.cu:24
    double eta = 0.0;         // integral of xi dt (not in the reference: for the NH conserved energy)
    double t = 0.0;
    double Lx0 = 0.0, Ly0 = 0.0;   // src/md_env.cu:30-31
    double last_P = 0.0;
    double p_A = 0.0;              // nph_area_momentum
    double gamma = 1.0;            // nph_aspect_ratio
    bool nph_init = false;
    double v_p = 0.0;              // nph_piston_velocity
    double F_p = 0.0;              // nph_piston_force
    bool v_p_init = false;
    bool piston_active = false;
};

inline Sim *S(orc_sim *s) { return reinterpret_cast<Sim *>(s); }
inline const Sim *S(const orc_sim *s) { return reinterpret_cast<const Sim *>(s); }

// include/md_cuda_common.hpp:20-29
inline double pbc_wrap(double x, double L) {
    if (L <= 0.0) return x;
    x = std::fmod(x, L);
    if (x < 0.0) x += L;
    return x;
}

inline double mass_of(const orc_params &p, int type) {
    return type == 0 ? p.mass_a : p.mass_b;   // src/md_cuda_common.cu:261
}

// Type-pair selection, src/md_cuda_common.cu:220-229: (0,0)->AA, (1,1)->BB,
// every other combination -> AB.
inline void select_pair(const orc_params &p, int ti, int tj, double &sig, double &eps) {
    if (ti == 0 && tj == 0) {
        sig = p.sigma_aa; eps = p.eps_aa;
    } else if (ti == 1 && tj == 1) {
        sig = p.sigma_bb; eps = p.eps_bb;
    } else {
        sig = p.sigma_ab; eps = p.eps_ab;
    }
}

// Energy shift Vc = eps * 4 (c^-12 - c^-6) from the dimensionless cutoff,
// src/md_cuda_common.cu:719-727.
inline double vc_const(double cutoff) {
    const double inv_c = 1.0 / cutoff;
    const double sr2 = inv_c * inv_c;
    const double sr6 = sr2 * sr2 * sr2;
    const double sr12 = sr6 * sr6;
    return 4.0 * (sr12 - sr6);
}

struct PairOut {
    double fx, fy, u, w, lam;
    bool in;
};

// One ordered pair (i <- j) with displacement d = r_i - r_j already minimum-
// imaged.  Pair math of src/md_cuda_common.cu:231-254 (force), :781-791
// (energy), :677-684 (virial), :349-356 (lambda), all in FP64.
inline PairOut pair_eval(const orc_params &p, double dx, double dy, int ti, int tj,
                         double vc_c) {
    PairOut o{0, 0, 0, 0, 0, false};
    double sig, eps;
    select_pair(p, ti, tj, sig, eps);
    const double rc = p.cutoff * sig;
    const double rc2 = rc * rc;
    const double r2 = dx * dx + dy * dy;
    if (r2 > 0.0 && r2 < rc2) {
        const double r2i = 1.0 / r2;
        const double sr2 = sig * sig * r2i;
        const double sr6 = sr2 * sr2 * sr2;
        const double sr12 = sr6 * sr6;
        const double tmp = 24.0 * eps * (2.0 * sr12 - sr6) * r2i;
        o.fx = tmp * dx;
        o.fy = tmp * dy;
        o.u = 4.0 * eps * (sr12 - sr6) - eps * vc_c;
        o.w = dx * o.fx + dy * o.fy;
        o.lam = o.fx * dx - o.fy * dy;
        o.in = true;
    }
    return o;
}

// Minimum image as the reference: d -= L * round(d / L) (src/md_cuda_common.cu:208-215).
inline double mic(double d, double L) { return d - L * std::round(d / L); }

struct CellGrid {
    int ncx = 3, ncy = 3;
};

inline double rc_max(const orc_params &p) {
    return p.cutoff * std::max(p.sigma_aa, std::max(p.sigma_bb, p.sigma_ab));
}

// Canonical cell grid: a periodic dimension gets max(3, floor(L / (s rc_max)))
// cells, a non-periodic one max(1, ...).
inline int grid_dim(double L, double side, bool periodic) {
    double q = std::floor(L / side);
    int lo = periodic ? 3 : 1;
    if (!(q >= lo)) return lo;
    if (q > 1.0e8) q = 1.0e8;
    return static_cast<int>(q);
}

inline int cell_coord(double x, double L, int n) {
    const double q = (x * static_cast<double>(n)) / L;
    if (!(q >= 0.0)) return 0;
    if (q >= static_cast<double>(n)) return n - 1;
    return static_cast<int>(q);
}

CellGrid make_grid(const Sim &s, bool periodic_y) {
    const double side = std::max(s.p.cell_scale, 1.0) * rc_max(s.p);
    CellGrid g;
    g.ncx = grid_dim(s.p.Lx, side, true);
    g.ncy = grid_dim(s.p.Ly, side, periodic_y);
    return g;
}

// Counting sort by cell; within a cell, ascending particle index.
void build_cells(const Sim &s, bool periodic_y, CellGrid &g, std::vector<uint32_t> &keys,
                 std::vector<uint32_t> &start, std::vector<uint32_t> &sorted) {
    g = make_grid(s, periodic_y);
    const size_t n = s.P.size();
    const size_t ncell = static_cast<size_t>(g.ncx) * g.ncy;
    keys.assign(n, 0);
    start.assign(ncell + 1, 0);
    sorted.assign(n, 0);
    for (size_t i = 0; i < n; ++i) {
        const int cx = cell_coord(s.P[i].px, s.p.Lx, g.ncx);
        const int cy = cell_coord(s.P[i].py, s.p.Ly, g.ncy);
        keys[i] = static_cast<uint32_t>(cy) * g.ncx + cx;
        start[keys[i] + 1]++;
    }
    for (size_t c = 0; c < ncell; ++c) start[c + 1] += start[c];
    std::vector<uint32_t> cur(start.begin(), start.end() - 1);
    for (size_t i = 0; i < n; ++i) sorted[cur[keys[i]]++] = static_cast<uint32_t>(i);
}

int nthreads_of(const orc_params &p) {
#ifdef _OPENMP
    return p.n_threads > 0 ? p.n_threads : omp_get_max_threads();
#else
    (void)p;
    return 1;
#endif
}

// Full pair pass: per-particle force (F, not F/m), energy, virial, lambda
// contributions (each ordered pair contributes its half to i).
void pair_pass(const Sim &s, bool periodic_y, std::vector<double> &fx, std::vector<double> &fy,
               std::vector<double> &ui, std::vector<double> &wi, std::vector<double> &li) {
    const int n = static_cast<int>(s.P.size());
    fx.assign(n, 0.0); fy.assign(n, 0.0); ui.assign(n, 0.0); wi.assign(n, 0.0); li.assign(n, 0.0);
    const orc_params &p = s.p;
    const double vcc = vc_const(p.cutoff);
    const double Lx = p.Lx, Ly = p.Ly;
    const int nt = nthreads_of(p);
    (void)nt;

    auto accumulate_pair = [&](int i, int j) {
        double dx = s.P[i].px - s.P[j].px;
        double dy = s.P[i].py - s.P[j].py;
        dx = mic(dx, Lx);
        if (periodic_y) dy = mic(dy, Ly);
        PairOut o = pair_eval(p, dx, dy, s.P[i].type, s.P[j].type, vcc);
        if (o.in) {
            fx[i] += o.fx; fy[i] += o.fy;
            ui[i] += 0.5 * o.u; wi[i] += 0.5 * o.w; li[i] += 0.5 * o.lam;
        }
    };

    if (p.force_mode == 0) {
#pragma omp parallel for schedule(static) num_threads(nt)
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j)
                if (j != i) accumulate_pair(i, j);
        return;
    }

    CellGrid g;
    std::vector<uint32_t> keys, start, sorted;
    build_cells(s, periodic_y, g, keys, start, sorted);
#pragma omp parallel for schedule(static) num_threads(nt)
    for (int i = 0; i < n; ++i) {
        const int cx = static_cast<int>(keys[i] % g.ncx);
        const int cy = static_cast<int>(keys[i] / g.ncx);
        for (int oy = -1; oy <= 1; ++oy) {
            int yy = cy + oy;
            if (periodic_y) {
                yy = (yy + g.ncy) % g.ncy;
            } else if (yy < 0 || yy >= g.ncy) {
                continue;
            }
            for (int ox = -1; ox <= 1; ++ox) {
                const int xx = (cx + ox + g.ncx) % g.ncx;
                const uint32_t c = static_cast<uint32_t>(yy) * g.ncx + xx;
                for (uint32_t k = start[c]; k < start[c + 1]; ++k) {
                    const int j = static_cast<int>(sorted[k]);
                    if (j != i) accumulate_pair(i, j);
                }
            }
        }
    }
}

// src/md_env.cu:1511-1542 (cal_forces) / :1544-1574 (cal_forces_piston, LJ part):
// acc = F / m.
void compute_forces(Sim &s, bool periodic_y) {
    std::vector<double> fx, fy, ui, wi, li;
    pair_pass(s, periodic_y, fx, fy, ui, wi, li);
    const int n = static_cast<int>(s.P.size());
    for (int i = 0; i < n; ++i) {
        const double m = mass_of(s.p, s.P[i].type);
        s.P[i].ax = fx[i] / m;
        s.P[i].ay = fy[i] / m;
    }
}

double total_K(const Sim &s) {   // src/md_cuda_common.cu:554-589
    double K = 0.0;
    for (const auto &q : s.P) K += 0.5 * mass_of(s.p, q.type) * (q.vx * q.vx + q.vy * q.vy);
    return K;
}

double sum_mv2(const Sim &s) {   // src/md_cuda_common.cu:608-611
    double a = 0.0;
    for (const auto &q : s.P) a += mass_of(s.p, q.type) * (q.vx * q.vx + q.vy * q.vy);
    return a;
}

struct WallObs {
    double upper = 0.0, energy = 0.0;
};

// src/md_cuda_common.cu:498-552 (piston_wall_kernel), src/md_env.cu:2148-2203.
WallObs piston_walls(Sim &s, bool update_acc) {
    WallObs w;
    const double k = s.p.k_piston;
    if (!(k > 0.0)) return w;
    const double Ly = s.p.Ly;
    for (auto &q : s.P) {
        double fy = 0.0;
        if (q.py > Ly) {
            const double ov = q.py - Ly;
            fy -= k * ov;
            w.upper += k * ov;
            w.energy += 0.5 * k * ov * ov;
        }
        if (q.py < 0.0) {
            const double ov = -q.py;
            fy += k * ov;
            w.energy += 0.5 * k * ov * ov;
        }
        if (update_acc && fy != 0.0) q.ay += fy / mass_of(s.p, q.type);
    }
    return w;
}

// src/md_env.cu:2205-2252: P = (sum m v^2 + 1/2 sum_ordered r.F) / (2 A).
double scalar_pressure(Sim &s, bool periodic_y) {
    std::vector<double> fx, fy, ui, wi, li;
    pair_pass(s, periodic_y, fx, fy, ui, wi, li);
    double W = 0.0;
    for (double v : wi) W += v;
    const double area = s.p.Lx * s.p.Ly;
    if (!(area > 0.0)) return 0.0;
    return (sum_mv2(s) + W) / (2.0 * area);
}

double total_U(Sim &s) {   // src/md_env.cu:2347-2360
    const bool py = !s.piston_active;
    std::vector<double> fx, fy, ui, wi, li;
    pair_pass(s, py, fx, fy, ui, wi, li);
    double U = 0.0;
    for (double v : ui) U += v;
    if (!s.piston_active) return U;
    return U + piston_walls(s, false).energy;
}

// src/md_env.cu:1457-1508 and src/md_cuda_common.cu:26-65.
void middle_reflect(Sim &s) {
    const double Lx = s.p.Lx, Ly = s.p.Ly, pdiv = s.p.devide_p;
    if (Lx <= 0.0 || Ly <= 0.0 || pdiv <= 0.0) return;
    const double smax = std::max(s.p.sigma_aa, std::max(s.p.sigma_bb, s.p.sigma_ab));
    double slab = pdiv * Lx + 2.0 * smax;
    if (slab > Lx) slab = Lx;
    const double center = 0.5 * Lx, x_min = center - 0.5 * slab, period = 2.0 * slab;
    for (auto &q : s.P) {
        const double y = pbc_wrap(q.py, Ly);
        double ym = std::fmod(q.px - x_min, period);
        if (ym < 0.0) ym += period;
        if (ym > slab) ym = period - ym;
        q.px = x_min + ym;
        q.py = y;
    }
}

// src/md_env.cu:1576-1637
void step_nve(Sim &s) {
    s.piston_active = false;
    const double dt = s.p.dt;
    for (auto &q : s.P) {   // step_half_vv_kernel, src/md_cuda_common.cu:381-394
        q.vx += 0.5 * dt * q.ax;
        q.vy += 0.5 * dt * q.ay;
        q.px += dt * q.vx;
        q.py += dt * q.vy;
        q.px = pbc_wrap(q.px, s.p.Lx);
        q.py = pbc_wrap(q.py, s.p.Ly);
    }
    compute_forces(s, true);
    for (auto &q : s.P) {   // step_2nd_half_vv_kernel, :443-452
        q.vx += 0.5 * dt * q.ax;
        q.vy += 0.5 * dt * q.ay;
    }
    s.t += dt;
}

// src/md_env.cu:1639-1728
void step_nvt(Sim &s, bool middle_wrap) {
    s.piston_active = false;
    const double dt = s.p.dt;
    const double g = 2.0 * static_cast<double>(s.P.size()) - 2.0;
    const double T0 = s.p.T_target, Q = s.p.Q;
    double K = total_K(s);
    s.xi += 0.5 * dt * (2.0 * K - g * T0) / Q;
    const double xi1 = s.xi;
    for (auto &q : s.P) {   // step_half_vv_nh_kernel, src/md_cuda_common.cu:454-478
        double vx = q.vx, vy = q.vy;
        vx += 0.5 * dt * (q.ax - xi1 * vx);
        vy += 0.5 * dt * (q.ay - xi1 * vy);
        q.vx = vx; q.vy = vy;
        q.px += dt * vx;
        q.py += dt * vy;
        q.px = pbc_wrap(q.px, s.p.Lx);
        q.py = pbc_wrap(q.py, s.p.Ly);
    }
    if (middle_wrap) middle_reflect(s);
    compute_forces(s, true);
    K = total_K(s);
    s.xi += 0.5 * dt * (2.0 * K - g * T0) / Q;
    const double xi2 = s.xi;
    for (auto &q : s.P) {   // step_2nd_half_vv_nh_kernel, :480-496
        double vx = q.vx, vy = q.vy;
        vx += 0.5 * dt * (q.ax - xi2 * vx);
        vy += 0.5 * dt * (q.ay - xi2 * vy);
        q.vx = vx; q.vy = vy;
    }
    s.eta += 0.5 * dt * (xi1 + xi2);
    s.t += dt;
}

void init_nph(Sim &s) {   // src/md_env.cu:1757-1778
    if (s.nph_init) return;
    s.gamma = s.p.Lx / s.p.Ly;
    s.p_A = std::max(s.p.barostat_mass, 1e-12) * s.p.barostat_area_rate_init;
    s.nph_init = true;
}

// src/md_env.cu:1795-1918.  Returns false where the reference aborts.
bool step_nph(Sim &s, double p_ext) {
    s.piston_active = false;
    init_nph(s);
    const double dt = s.p.dt;
    const double W = std::max(s.p.barostat_mass, 1e-12);
    const double Lx_old = s.p.Lx, Ly_old = s.p.Ly;
    const double area_old = Lx_old * Ly_old;
    if (!(area_old > 0.0)) return false;
    const double P_old = scalar_pressure(s, true);
    for (auto &q : s.P) {   // step_half_nph_velocity_kernel, :411-420
        q.vx += 0.5 * dt * q.ax;
        q.vy += 0.5 * dt * q.ay;
    }
    s.p_A += 0.5 * dt * (P_old - p_ext);
    double area_new = area_old + dt * (s.p_A / W);
    if (!(area_new > 0.0) || !std::isfinite(area_new)) return false;
    if (!(s.gamma > 0.0) || !std::isfinite(s.gamma)) s.gamma = Lx_old / Ly_old;
    double Ly_new = std::sqrt(area_new / s.gamma);
    double Lx_new = s.gamma * Ly_new;
    const double floor_ = std::max(1e-12, s.p.barostat_height_min_ratio * s.Ly0);
    const double ceil_ = std::max(floor_, s.p.barostat_height_max_ratio * s.Ly0);
    bool clamped = false;
    if (Ly_new < floor_) {
        Ly_new = floor_; Lx_new = s.gamma * Ly_new; area_new = Lx_new * Ly_new; clamped = true;
    } else if (Ly_new > ceil_) {
        Ly_new = ceil_; Lx_new = s.gamma * Ly_new; area_new = Lx_new * Ly_new; clamped = true;
    }
    const double lin = std::sqrt(area_new / area_old);
    const double vsc = 1.0 / lin;
    for (auto &q : s.P) {   // step_nph_scale_and_drift_kernel, :422-440
        q.vx *= vsc; q.vy *= vsc;
        q.px = lin * q.px + dt * q.vx;
        q.py = lin * q.py + dt * q.vy;
        q.px = pbc_wrap(q.px, Lx_new);
        q.py = pbc_wrap(q.py, Ly_new);
    }
    s.p.Lx = Lx_new;
    s.p.Ly = Ly_new;
    compute_forces(s, true);
    const double P_new = scalar_pressure(s, true);
    for (auto &q : s.P) {
        q.vx += 0.5 * dt * q.ax;
        q.vy += 0.5 * dt * q.ay;
    }
    s.p_A += 0.5 * dt * (P_new - p_ext);
    if (clamped) {
        if (s.p.Ly <= floor_ && s.p_A < 0.0) s.p_A = 0.0;
        if (s.p.Ly >= ceil_ && s.p_A > 0.0) s.p_A = 0.0;
    }
    s.last_P = P_new;
    s.t += dt;
    return true;
}

void cal_forces_piston(Sim &s) {   // src/md_env.cu:1544-1574
    compute_forces(s, false);
    s.F_p = piston_walls(s, true).upper;
}

void init_piston(Sim &s) {   // src/md_env.cu:1780-1793
    if (!s.v_p_init) {
        s.v_p = 0.0;
        s.v_p_init = true;
    }
    if (!s.piston_active) {
        s.piston_active = true;
        cal_forces_piston(s);
        s.last_P = (s.p.Lx > 0.0) ? (s.F_p / s.p.Lx) : 0.0;
    }
}

// src/md_env.cu:1920-2023
void step_piston(Sim &s, double p_ext) {
    init_piston(s);
    const double dt = s.p.dt;
    const double W = std::max(s.p.barostat_mass, 1e-12);
    const double Lx = s.p.Lx, Ly_old = s.p.Ly;
    const double acc_old = (s.F_p - p_ext * Lx) / W;
    for (auto &q : s.P) {   // step_half_vv_piston_kernel, :396-409
        q.vx += 0.5 * dt * q.ax;
        q.vy += 0.5 * dt * q.ay;
        q.px += dt * q.vx;
        q.py += dt * q.vy;
        q.px = pbc_wrap(q.px, Lx);
    }
    s.v_p += 0.5 * dt * acc_old;
    double Ly_new = Ly_old + dt * s.v_p;
    const double floor_ = std::max(1e-12, s.p.barostat_height_min_ratio * s.Ly0);
    const double ceil_ = std::max(floor_, s.p.barostat_height_max_ratio * s.Ly0);
    bool clamped = false;
    if (Ly_new < floor_) {
        Ly_new = floor_; clamped = true;
        if (s.v_p < 0.0) s.v_p = 0.0;
    } else if (Ly_new > ceil_) {
        Ly_new = ceil_; clamped = true;
        if (s.v_p > 0.0) s.v_p = 0.0;
    }
    s.p.Ly = Ly_new;
    cal_forces_piston(s);
    for (auto &q : s.P) {
        q.vx += 0.5 * dt * q.ax;
        q.vy += 0.5 * dt * q.ay;
    }
    const double acc_new = (s.F_p - p_ext * Lx) / W;
    s.v_p += 0.5 * dt * acc_new;
    if (clamped) {
        if (s.p.Ly <= floor_ && s.v_p < 0.0) s.v_p = 0.0;
        if (s.p.Ly >= ceil_ && s.v_p > 0.0) s.v_p = 0.0;
    }
    s.last_P = s.F_p / Lx;
    s.t += dt;
}

}  // namespace

extern "C" {

double orc_pbc_wrap(double x, double L) { return pbc_wrap(x, L); }

void orc_pair(double dx, double dy, int32_t ti, int32_t tj, const orc_params *p, double *fx,
              double *fy, double *u, int32_t *in_cut) {
    PairOut o = pair_eval(*p, dx, dy, ti, tj, vc_const(p->cutoff));
    if (fx) *fx = o.fx;
    if (fy) *fy = o.fy;
    if (u) *u = o.u;
    if (in_cut) *in_cut = o.in ? 1 : 0;
}

orc_sim *orc_create(const orc_params *p, const orc_particle *frame, int64_t n) {
    Sim *s = new Sim();
    s->p = *p;
    s->P.assign(frame, frame + n);
    s->Lx0 = p->Lx;
    s->Ly0 = p->Ly;
    return reinterpret_cast<orc_sim *>(s);
}

void orc_destroy(orc_sim *s) { delete S(s); }

void orc_get_frame(const orc_sim *s, orc_particle *out) {
    std::memcpy(out, S(s)->P.data(), S(s)->P.size() * sizeof(orc_particle));
}

void orc_set_frame(orc_sim *s, const orc_particle *frame) {
    std::memcpy(S(s)->P.data(), frame, S(s)->P.size() * sizeof(orc_particle));
}

void orc_set_force_mode(orc_sim *s, int32_t mode) { S(s)->p.force_mode = mode; }

void orc_compute_forces(orc_sim *s, int32_t periodic_y) { compute_forces(*S(s), periodic_y != 0); }

void orc_pair_pass(const orc_sim *s, int32_t periodic_y, double *acc_xy, double *u_i, double *w_i,
                   double *lam_i, double eps_lambda) {
    const Sim &sim = *S(s);
    std::vector<double> fx, fy, ui, wi, li;
    pair_pass(sim, periodic_y != 0, fx, fy, ui, wi, li);
    const size_t n = sim.P.size();
    for (size_t i = 0; i < n; ++i) {
        const double m = mass_of(sim.p, sim.P[i].type);
        if (acc_xy) {
            acc_xy[2 * i] = fx[i] / m;
            acc_xy[2 * i + 1] = fy[i] / m;
        }
        if (u_i) u_i[i] = ui[i];
        if (w_i) w_i[i] = wi[i];
        if (lam_i) lam_i[i] = -eps_lambda * li[i];
    }
}

double orc_total_K(const orc_sim *s) { return total_K(*S(s)); }

double orc_total_U(orc_sim *s) { return total_U(*S(s)); }

// src/md_env.cu:2254-2266
double orc_instant_pressure(orc_sim *s) {
    Sim &sim = *S(s);
    if (sim.piston_active) {
        sim.F_p = piston_walls(sim, false).upper;
        sim.last_P = (sim.p.Lx > 0.0) ? sim.F_p / sim.p.Lx : 0.0;
        return sim.last_P;
    }
    sim.last_P = scalar_pressure(sim, true);
    return sim.last_P;
}

// Pressure-tensor profile in nb slabs along y: local_pressure_tensor_profile_kernel
// (src/md_cuda_common.cu:879-1030) driven by get_pressure_profile
// (src/md_env.cu:3821-3918) with one rank.  Particle i adds m v_a v_b to the
// slab of its y; each ordered pair inside the cutoff adds 1/2 d_a F_b to the
// slab of the midpoint y_i - d_y / 2; slab = int(y / Ly * nb) clamped to
// [0, nb) after wrapping y (periodic) or clamping it into [0, Ly).  Positions
// are wrapped first, displacements take the minimum image.  Each contribution
// is rounded to int64 units of 2^-32 and the units are summed as integers (the
// library's definition, which makes the sums independent of their order; the
// reference adds doubles with atomics).  out: Pxx[nb], Pyy[nb], Pxy[nb], each
// slab's sum divided by its area Lx Ly / nb.
void orc_pressure_profile(orc_sim *s, int32_t nb, int32_t periodic_y, double *out) {
    Sim &sim = *S(s);
    const orc_params &p = sim.p;
    const double Lx = p.Lx, Ly = p.Ly;
    const int n = static_cast<int>(sim.P.size());
    const bool py = periodic_y != 0;
    std::vector<long long> acc(static_cast<size_t>(3 * nb), 0);
    auto fixed = [](double v) {
        double q = std::rint(v * 4294967296.0);
        q = q > 4.6e18 ? 4.6e18 : (q < -4.6e18 ? -4.6e18 : q);
        return static_cast<long long>(q);
    };
    auto bin = [&](double y) {
        if (py) {
            y = pbc_wrap(y, Ly);
        } else if (y < 0.0) {
            y = 0.0;
        } else if (y >= Ly && Ly > 0.0) {
            y = Ly * (1.0 - 1e-12);
        }
        const double q = Ly > 0.0 ? y / Ly : 0.0;
        int b = static_cast<int>(q * static_cast<double>(nb));
        return b < 0 ? 0 : (b >= nb ? nb - 1 : b);
    };
    std::vector<double> xw(n), yw(n);
    for (int i = 0; i < n; ++i) {
        xw[i] = pbc_wrap(sim.P[i].px, Lx);
        yw[i] = py ? pbc_wrap(sim.P[i].py, Ly) : sim.P[i].py;
        const double m = mass_of(p, sim.P[i].type);
        const int b = bin(yw[i]);
        acc[b] += fixed(m * sim.P[i].vx * sim.P[i].vx);
        acc[nb + b] += fixed(m * sim.P[i].vy * sim.P[i].vy);
        acc[2 * nb + b] += fixed(m * sim.P[i].vx * sim.P[i].vy);
    }
    const double vcc = vc_const(p.cutoff);
    auto pair = [&](int i, int j) {
        double dx = mic(xw[i] - xw[j], Lx);
        double dy = yw[i] - yw[j];
        if (py) dy = mic(dy, Ly);
        const PairOut o = pair_eval(p, dx, dy, sim.P[i].type, sim.P[j].type, vcc);
        if (!o.in) return;
        const int b = bin(yw[i] - 0.5 * dy);
        acc[b] += fixed(0.5 * dx * o.fx);
        acc[nb + b] += fixed(0.5 * dy * o.fy);
        acc[2 * nb + b] += fixed(0.5 * dx * o.fy);
    };
    if (p.force_mode == 0) {
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j)
                if (j != i) pair(i, j);
    } else {
        CellGrid g;
        std::vector<uint32_t> keys, start, sorted;
        build_cells(sim, py, g, keys, start, sorted);
        for (int i = 0; i < n; ++i) {
            const int cx = static_cast<int>(keys[i] % g.ncx);
            const int cy = static_cast<int>(keys[i] / g.ncx);
            for (int oy = -1; oy <= 1; ++oy) {
                int yy = cy + oy;
                if (py) {
                    yy = (yy + g.ncy) % g.ncy;
                } else if (yy < 0 || yy >= g.ncy) {
                    continue;
                }
                for (int ox = -1; ox <= 1; ++ox) {
                    const int xx = (cx + ox + g.ncx) % g.ncx;
                    const uint32_t c = static_cast<uint32_t>(yy) * g.ncx + xx;
                    for (uint32_t k = start[c]; k < start[c + 1]; ++k) {
                        const int j = static_cast<int>(sorted[k]);
                        if (j != i) pair(i, j);
                    }
                }
            }
        }
    }
    const double area = Lx * Ly / static_cast<double>(nb);
    for (int k = 0; k < 3 * nb; ++k) {
        const double v = static_cast<double>(acc[k]) / 4294967296.0;
        out[k] = area > 0.0 ? v / area : v;
    }
}

// src/md_env.cu:2362-2406: -eps * 1/2 sum_ordered (Fx dx - Fy dy).
double orc_partial_U_lambda(orc_sim *s, double eps_lambda) {
    Sim &sim = *S(s);
    std::vector<double> fx, fy, ui, wi, li;
    pair_pass(sim, !sim.piston_active, fx, fy, ui, wi, li);
    double acc = 0.0;
    for (double v : li) acc += -eps_lambda * v;
    return acc;
}

void orc_get_obs(const orc_sim *s, orc_obs *o) {
    const Sim &sim = *S(s);
    std::memset(o, 0, sizeof(*o));
    o->K = total_K(sim);
    o->xi = sim.xi;
    o->p_A = sim.p_A;
    o->v_p = sim.v_p;
    o->F_p = sim.F_p;
    o->Lx = sim.p.Lx;
    o->Ly = sim.p.Ly;
    o->t = sim.t;
    o->last_P = sim.last_P;
    o->area_rate = sim.p_A / std::max(sim.p.barostat_mass, 1e-12);
    o->eta = sim.eta;
}

void orc_step_nve(orc_sim *s, int32_t nsteps) {
    for (int k = 0; k < nsteps; ++k) step_nve(*S(s));
}

void orc_step_nvt(orc_sim *s, int32_t nsteps, int32_t middle_wrap) {
    for (int k = 0; k < nsteps; ++k) step_nvt(*S(s), middle_wrap != 0);
}

void orc_middle_reflect(orc_sim *s) { middle_reflect(*S(s)); }

void orc_step_nph(orc_sim *s, int32_t nsteps, double p_ext) {
    for (int k = 0; k < nsteps; ++k)
        if (!step_nph(*S(s), p_ext)) return;
}

void orc_step_piston(orc_sim *s, int32_t nsteps, double p_ext) {
    for (int k = 0; k < nsteps; ++k) step_piston(*S(s), p_ext);
}

// src/md_energy_minimizer.cu:122-272 (single domain; halos are empty there too).
void orc_fire(orc_sim *sp, const orc_fire_params *fp, double force_tol, int32_t max_steps,
              orc_fire_result *res, int32_t *energy_steps, double *energy_trace, int32_t cap) {
    Sim &s = *S(sp);
    s.piston_active = false;
    for (auto &q : s.P) {   // initialize_from_frame, :343-380
        q.px = pbc_wrap(q.px, s.p.Lx);
        q.py = pbc_wrap(q.py, s.p.Ly);
        q.vx = q.vy = q.ax = q.ay = 0.0;
    }
    int ns = 0;
    auto push = [&](int step, double U) {
        if (ns < cap) {
            energy_steps[ns] = step;
            energy_trace[ns] = U;
        }
        ++ns;
    };
    auto potential = [&]() {
        std::vector<double> fx, fy, ui, wi, li;
        pair_pass(s, true, fx, fy, ui, wi, li);
        double U = 0.0;
        for (double v : ui) U += v;
        return U;
    };
    double dt = fp->dt_init, alpha = fp->alpha0;
    int streak = 0;
    const int interval = fp->energy_interval > 0 ? fp->energy_interval : 1000;
    double last_U = potential();
    push(0, last_U);
    bool converged = false;
    int steps = 0;
    double max_force = 0.0;
    for (int step = 0; step < max_steps; ++step) {
        compute_forces(s, true);
        for (auto &q : s.P) {   // velocity_kick_kernel, :24-30
            q.vx += dt * q.ax;
            q.vy += dt * q.ay;
        }
        // reduce_force_metrics, :442-475
        double mf = 0.0, power = 0.0;
        for (const auto &q : s.P) {
            mf = std::max(mf, std::sqrt(q.ax * q.ax + q.ay * q.ay));
            power += q.vx * q.ax + q.vy * q.ay;
        }
        max_force = mf;
        if (mf < force_tol) {
            converged = true;
            steps = step;
            last_U = potential();
            push(step + 1, last_U);
            break;
        }
        if (power > 0.0) {
            streak += 1;
            if (streak > fp->N_min) {
                dt = std::min(dt * fp->f_inc, fp->dt_max);
                alpha = std::max(0.0, alpha * 0.99);
            }
        } else {
            dt *= fp->f_dec;
            streak = 0;
            alpha = fp->alpha0;
            for (auto &q : s.P) q.vx = q.vy = 0.0;   // zero_velocity_kernel, :32-38
        }
        for (auto &q : s.P) {   // fire_velocity_mix_kernel, :40-60 (per particle)
            const double vmag = std::sqrt(q.vx * q.vx + q.vy * q.vy);
            const double fmag = std::sqrt(q.ax * q.ax + q.ay * q.ay);
            if (vmag > 0.0 && fmag > 0.0) {
                const double sc = vmag / fmag;
                const double nvx = (1.0 - alpha) * q.vx + alpha * sc * q.ax;
                const double nvy = (1.0 - alpha) * q.vy + alpha * sc * q.ay;
                q.vx = nvx;
                q.vy = nvy;
            }
        }
        for (auto &q : s.P) {   // drift_kernel, :62-71 (+ optional displacement cap)
            double ddx = dt * q.vx, ddy = dt * q.vy;
            if (fp->max_step > 0.0) {
                const double d = std::sqrt(ddx * ddx + ddy * ddy);
                if (d > fp->max_step) {
                    const double f = fp->max_step / d;
                    ddx *= f;
                    ddy *= f;
                }
            }
            q.px = pbc_wrap(q.px + ddx, s.p.Lx);
            q.py = pbc_wrap(q.py + ddy, s.p.Ly);
        }
        if ((step + 1) % interval == 0) {
            last_U = potential();
            push(step + 1, last_U);
        }
    }
    if (!converged) {
        steps = max_steps;
        int last_step = (ns > 0 && ns <= cap) ? energy_steps[ns - 1] : -1;
        if (ns == 0 || last_step != max_steps) {
            last_U = potential();
            push(max_steps, last_U);
        }
    }
    res->steps = steps;
    res->converged = converged ? 1 : 0;
    res->max_force = max_force;
    res->potential_energy = potential();   // :270
    res->n_energy = ns;
}

int64_t orc_cell_list(const orc_sim *s, int32_t periodic_y, int32_t *ncx, int32_t *ncy,
                      uint32_t *keys, uint32_t *cell_start, uint32_t *sorted) {
    CellGrid g;
    std::vector<uint32_t> k, st, so;
    build_cells(*S(s), periodic_y != 0, g, k, st, so);
    if (ncx) *ncx = g.ncx;
    if (ncy) *ncy = g.ncy;
    if (keys) std::memcpy(keys, k.data(), k.size() * 4);
    if (cell_start) std::memcpy(cell_start, st.data(), st.size() * 4);
    if (sorted) std::memcpy(sorted, so.data(), so.size() * 4);
    return static_cast<int64_t>(g.ncx) * g.ncy;
}

}  // extern "C"
