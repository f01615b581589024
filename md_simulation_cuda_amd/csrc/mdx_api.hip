// Host side of libmdx: the simulation handle, kernel launch sequences per
// ensemble, and the extern "C" entry points declared in include/mdx.h.
//
// Step sequencing restates src/md_env.cu:1576-2023 (NVE, Nose-Hoover, NPH,
// piston) and src/md_energy_minimizer.cu:122-272 (FIRE) for one periodic
// domain: the MPI slab exchange (src/md_env.cu:757-1274) is replaced by a cell
// list rebuilt on the device every step.
#include "mdx.h"
#include "mdx_kernels.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define MDX_HIP(...)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (__VA_ARGS__);                                                       \
        if (e_ != hipSuccess)                                                                \
            return fail(MDX_ERR_HIP, std::string(#__VA_ARGS__) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define MDX_TRY(...)                   \
    do {                               \
        int rc_ = (__VA_ARGS__);       \
        if (rc_ != MDX_OK) return rc_; \
    } while (0)

using namespace mdx;

int blocks_for(int64_t n) { return static_cast<int>((n + kBlock - 1) / kBlock); }

}  // namespace

struct mdx_sim {
    mdx_params p{};
    int64_t n = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    PairTable pt{};
    double rc_max = 0.0;
    int prec = MDX_PREC_COMP32;

    // device buffers
    double2 *x = nullptr, *v = nullptr;
    void *a = nullptr;      // n * 16 bytes (float2 or double2 view)
    uint8_t *cls = nullptr;
    uint32_t *key = nullptr, *sorted = nullptr, *scell = nullptr;
    void *img = nullptr;    // n * 16 bytes
    uint8_t *icls = nullptr;
    double *partials = nullptr;
    DevScalars *sc = nullptr;
    uint32_t *count = nullptr, *cs = nullptr, *bsum = nullptr;
    int64_t ncell_cap = 0;

    // host mirrors
    std::vector<int32_t> type;
    std::vector<double> theta;
    double t = 0.0;
    double Lx = 0.0, Ly = 0.0, Lx0 = 0.0, Ly0 = 0.0;
    Grid grid{};
    bool cells_valid = false;
    int cells_py = 1;
    bool K_valid = false, W_valid = false;
    bool piston_active = false, nph_init = false, vp_init = false;

    // timing
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> force_ev, step_ev;
    double force_ms = 0.0, step_ms = 0.0;
    int64_t force_n = 0, step_n = 0;
};

namespace {

void make_pair_table(const mdx_params &p, PairTable &pt, double &rc_max) {
    const double sig[3] = {p.sigma_aa, p.sigma_bb, p.sigma_ab};
    const double eps[3] = {p.eps_aa, p.eps_bb, p.eps_ab};
    const double inv_c = 1.0 / p.cutoff;
    const double s2 = inv_c * inv_c, s6 = s2 * s2 * s2, s12 = s6 * s6;
    const double vcc = 4.0 * (s12 - s6);   // src/md_cuda_common.cu:719-727
    const float cutf = static_cast<float>(p.cutoff);
    for (int k = 0; k < 3; ++k) {
        const double rc = p.cutoff * sig[k];
        pt.rc2[k] = rc * rc;
        pt.sig2[k] = sig[k] * sig[k];
        pt.eps24[k] = 24.0 * eps[k];
        pt.eps4[k] = 4.0 * eps[k];
        pt.vc[k] = eps[k] * vcc;
        // FP32 constants as src/md_cuda_common.cu:220-241 forms them
        const float sf = static_cast<float>(sig[k]), ef = static_cast<float>(eps[k]);
        const float rcf = cutf * sf;
        pt.rc2f[k] = rcf * rcf;
        pt.sig2f[k] = sf * sf;
        pt.eps24f[k] = 24.0f * ef;
        pt.eps4f[k] = 4.0f * ef;
    }
    pt.m[0] = p.mass_a;
    pt.m[1] = p.mass_b;
    pt.m[2] = p.mass_b;
    for (int k = 0; k < 3; ++k) pt.inv_m[k] = 1.0 / pt.m[k];
    // FP32 (sigma/r)^12 / r^2 overflows below r ~ 2.3e-3 sigma; evaluate such
    // overlapping pairs in FP64.
    const double smin = std::min(sig[0], std::min(sig[1], sig[2]));
    pt.r2_guard = static_cast<float>(1.0e-4 * smin * smin);
    rc_max = p.cutoff * std::max(sig[0], std::max(sig[1], sig[2]));
}

// Canonical grid: identical expressions in the oracle.
int grid_dim(double L, double side, bool periodic) {
    double q = std::floor(L / side);
    const int lo = periodic ? 3 : 1;
    if (!(q >= lo)) return lo;
    if (q > 1.0e8) q = 1.0e8;
    return static_cast<int>(q);
}

Grid grid_for(const mdx_sim &s, double Lx, double Ly, int periodic_y) {
    Grid g{};
    const double side = std::max(s.p.cell_scale, 1.0) * s.rc_max;
    g.Lx = Lx;
    g.Ly = Ly;
    g.ncx = grid_dim(Lx, side, true);
    g.ncy = grid_dim(Ly, side, periodic_y != 0);
    g.w = Lx / g.ncx;
    g.h = Ly / g.ncy;
    g.Lxf = static_cast<float>(Lx);
    g.Lyf = static_cast<float>(Ly);
    g.wf = static_cast<float>(g.w);
    g.hf = static_cast<float>(g.h);
    g.periodic_y = periodic_y;
    g.small_x = g.w < s.rc_max ? 1 : 0;
    g.small_y = g.h < s.rc_max ? 1 : 0;
    return g;
}

int ensure_cells(mdx_sim &s, const Grid &g) {
    const int64_t ncell = static_cast<int64_t>(g.ncx) * g.ncy;
    if (ncell > static_cast<int64_t>(UINT32_MAX) / 2) return fail(MDX_ERR_CAPACITY, "cell grid too large");
    if (ncell <= s.ncell_cap) return MDX_OK;
    const int64_t cap = std::max<int64_t>(ncell + ncell / 4, 1024);
    MDX_HIP(hipStreamSynchronize(s.stream));
    if (s.count) MDX_HIP(hipFree(s.count));
    if (s.cs) MDX_HIP(hipFree(s.cs));
    if (s.bsum) MDX_HIP(hipFree(s.bsum));
    s.count = s.cs = s.bsum = nullptr;
    MDX_HIP(hipMalloc(&s.count, cap * sizeof(uint32_t)));
    MDX_HIP(hipMalloc(&s.cs, (cap + 1) * sizeof(uint32_t)));
    MDX_HIP(hipMalloc(&s.bsum, ((cap + kScanChunk - 1) / kScanChunk) * sizeof(uint32_t)));
    MDX_HIP(hipMemset(s.count, 0, cap * sizeof(uint32_t)));
    s.ncell_cap = cap;
    return MDX_OK;
}

DriftArgs drift_args(const mdx_sim &s, uint32_t flags) {
    DriftArgs d{};
    d.dt = s.p.dt;
    d.T0 = s.p.T_target;
    d.Q = s.p.Q;
    d.g_dof = 2.0 * static_cast<double>(s.n) - 2.0;   // src/md_env.cu:1667
    d.middle_wrap = 0;
    if ((flags & MDX_STEP_MIDDLE_WRAP) && s.Lx > 0.0 && s.Ly > 0.0 && s.p.devide_p > 0.0) {
        // src/md_env.cu:1466-1473
        const double smax = std::max(s.p.sigma_aa, std::max(s.p.sigma_bb, s.p.sigma_ab));
        double slab = s.p.devide_p * s.Lx + 2.0 * smax;
        if (slab > s.Lx) slab = s.Lx;
        d.middle_wrap = 1;
        d.slab_width = slab;
        d.slab_xmin = 0.5 * s.Lx - 0.5 * slab;
        d.slab_period = 2.0 * slab;
    }
    return d;
}

PairArgs pair_args(const mdx_sim &s, double p_ext) {
    PairArgs a{};
    a.dt = s.p.dt;
    a.p_ext = p_ext;
    a.W_baro = std::max(s.p.barostat_mass, 1e-12);
    a.k_piston = s.p.k_piston;
    return a;
}

template <int KIND>
int launch_drift(mdx_sim &s, const Grid &g, const DriftArgs &da) {
    const int nb = blocks_for(s.n);
    const int n = static_cast<int>(s.n);
    if (s.prec == MDX_PREC_COMP32)
        k_drift<KIND, float2><<<nb, kBlock, 0, s.stream>>>(n, s.x, s.v, static_cast<const float2 *>(s.a), s.cls,
                                                            s.key, s.count, g, da, s.pt, s.sc, s.partials);
    else
        k_drift<KIND, double2><<<nb, kBlock, 0, s.stream>>>(n, s.x, s.v, static_cast<const double2 *>(s.a), s.cls,
                                                             s.key, s.count, g, da, s.pt, s.sc, s.partials);
    MDX_HIP(hipGetLastError());
    return MDX_OK;
}

// scan -> scatter -> per-cell sort + image (counts must be in place).
int launch_cells(mdx_sim &s, const Grid &g, int wrap, const int *skip) {
    const int ncell = g.ncx * g.ncy;
    const int nchunk = (ncell + kScanChunk - 1) / kScanChunk;
    k_scan_a<<<nchunk, kBlock, 0, s.stream>>>(s.count, ncell, s.bsum, skip);
    k_scan_b<<<nchunk, kBlock, 0, s.stream>>>(s.count, ncell, s.bsum, s.cs, skip);
    k_scatter<<<blocks_for(s.n), kBlock, 0, s.stream>>>(static_cast<int>(s.n), s.key, s.count, s.cs, s.sorted, skip);
    const int nbc = blocks_for(ncell);
    if (s.prec == MDX_PREC_FP64)
        k_cell_image<double2><<<nbc, kBlock, 0, s.stream>>>(ncell, s.cs, s.sorted, s.x, s.cls, g, wrap,
                                                            static_cast<double2 *>(s.img), s.icls, s.scell, skip);
    else
        k_cell_image<float2><<<nbc, kBlock, 0, s.stream>>>(ncell, s.cs, s.sorted, s.x, s.cls, g, wrap,
                                                           static_cast<float2 *>(s.img), s.icls, s.scell, skip);
    MDX_HIP(hipGetLastError());
    return MDX_OK;
}

// Rebuild the cell list from the current positions (wrapped for the keys and
// the image, master positions untouched).
int rebuild_cells(mdx_sim &s, int periodic_y) {
    const Grid g = grid_for(s, s.Lx, s.Ly, periodic_y);
    MDX_TRY(ensure_cells(s, g));
    DriftArgs da{};
    MDX_TRY(launch_drift<DK_KEYS>(s, g, da));
    MDX_TRY(launch_cells(s, g, 1, nullptr));
    s.grid = g;
    s.cells_valid = true;
    s.cells_py = periodic_y;
    return MDX_OK;
}

int ensure_cells_for(mdx_sim &s, int periodic_y) {
    if (s.cells_valid && s.cells_py == periodic_y) return MDX_OK;
    return rebuild_cells(s, periodic_y);
}

void timing_begin(mdx_sim &s, std::vector<std::pair<hipEvent_t, hipEvent_t>> &v) {
    if (!s.timing) return;
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return;
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return;
    }
    (void)hipEventRecord(a, s.stream);
    v.emplace_back(a, b);
}

void timing_end(mdx_sim &s, std::vector<std::pair<hipEvent_t, hipEvent_t>> &v) {
    if (!s.timing || v.empty()) return;
    (void)hipEventRecord(v.back().second, s.stream);
}

template <int MODE, bool OBS, bool WALLS>
int launch_pair(mdx_sim &s, const Grid &g, const PairArgs &pa) {
    const int nb = blocks_for(s.n);
    const int n = static_cast<int>(s.n);
    timing_begin(s, s.force_ev);
    switch (s.prec) {
    case MDX_PREC_COMP32:
        k_pair<float2, ACC_COMP32, MODE, OBS, WALLS><<<nb, kBlock, 0, s.stream>>>(
            n, s.cs, static_cast<const float2 *>(s.img), s.icls, s.scell, s.sorted, g, s.pt, pa,
            static_cast<float2 *>(s.a), s.v, s.x, s.sc, s.partials);
        break;
    case MDX_PREC_ACC64:
        k_pair<float2, ACC_FP64, MODE, OBS, WALLS><<<nb, kBlock, 0, s.stream>>>(
            n, s.cs, static_cast<const float2 *>(s.img), s.icls, s.scell, s.sorted, g, s.pt, pa,
            static_cast<double2 *>(s.a), s.v, s.x, s.sc, s.partials);
        break;
    default:
        k_pair<double2, ACC_FP64, MODE, OBS, WALLS><<<nb, kBlock, 0, s.stream>>>(
            n, s.cs, static_cast<const double2 *>(s.img), s.icls, s.scell, s.sorted, g, s.pt, pa,
            static_cast<double2 *>(s.a), s.v, s.x, s.sc, s.partials);
        break;
    }
    timing_end(s, s.force_ev);
    MDX_HIP(hipGetLastError());
    return MDX_OK;
}

int read_scalars(mdx_sim &s, DevScalars &h) {
    MDX_HIP(hipMemcpyAsync(&h, s.sc, sizeof(DevScalars), hipMemcpyDeviceToHost, s.stream));
    MDX_HIP(hipStreamSynchronize(s.stream));
    return MDX_OK;
}

template <typename T>
int write_scalar(mdx_sim &s, T DevScalars::*field, T value) {
    T *dst = &(s.sc->*field);
    MDX_HIP(hipMemcpyAsync(dst, &value, sizeof(T), hipMemcpyHostToDevice, s.stream));
    MDX_HIP(hipStreamSynchronize(s.stream));
    return MDX_OK;
}

int kinetic(mdx_sim &s) {
    k_kinetic<<<blocks_for(s.n), kBlock, 0, s.stream>>>(static_cast<int>(s.n), s.v, s.cls, s.pt, s.sc, s.partials);
    MDX_HIP(hipGetLastError());
    s.K_valid = true;
    return MDX_OK;
}

// cal_forces (src/md_env.cu:1511-1542): accelerations at the current positions.
int compute_forces(mdx_sim &s, int periodic_y) {
    MDX_TRY(rebuild_cells(s, periodic_y));
    PairArgs pa = pair_args(s, 0.0);
    MDX_TRY(launch_pair<FM_PLAIN, false, false>(s, s.grid, pa));
    s.W_valid = false;
    return MDX_OK;
}

// cal_forces_piston (src/md_env.cu:1544-1574): LJ with open y plus the walls.
int compute_forces_piston(mdx_sim &s) {
    MDX_TRY(rebuild_cells(s, 0));
    PairArgs pa = pair_args(s, 0.0);
    MDX_TRY(launch_pair<FM_PLAIN, false, true>(s, s.grid, pa));
    s.W_valid = false;
    return MDX_OK;
}

int check_n(const mdx_sim &s) {
    if (s.n <= 0) return fail(MDX_ERR_STATE, "empty system");
    return MDX_OK;
}

int step_nve(mdx_sim &s) {   // src/md_env.cu:1576-1637
    s.piston_active = false;
    const Grid g = grid_for(s, s.Lx, s.Ly, 1);
    MDX_TRY(ensure_cells(s, g));
    MDX_TRY(launch_drift<DK_NVE>(s, g, drift_args(s, 0)));
    MDX_TRY(launch_cells(s, g, 0, nullptr));
    MDX_TRY(launch_pair<FM_NVE, false, false>(s, g, pair_args(s, 0.0)));
    s.grid = g;
    s.cells_valid = true;
    s.cells_py = 1;
    s.K_valid = false;
    s.W_valid = false;
    s.t += s.p.dt;
    return MDX_OK;
}

int step_nvt(mdx_sim &s, uint32_t flags) {   // src/md_env.cu:1639-1728
    s.piston_active = false;
    if (!s.K_valid) MDX_TRY(kinetic(s));
    const Grid g = grid_for(s, s.Lx, s.Ly, 1);
    MDX_TRY(ensure_cells(s, g));
    MDX_TRY(launch_drift<DK_NVT>(s, g, drift_args(s, flags)));
    MDX_TRY(launch_cells(s, g, 0, nullptr));
    MDX_TRY(launch_pair<FM_NVT, false, false>(s, g, pair_args(s, 0.0)));
    s.grid = g;
    s.cells_valid = true;
    s.cells_py = 1;
    s.K_valid = true;
    s.W_valid = false;
    s.t += s.p.dt;
    return MDX_OK;
}

int step_nph(mdx_sim &s, double p_ext) {   // src/md_env.cu:1795-1918
    s.piston_active = false;
    if (!s.nph_init) {   // init_nph_state_if_needed, src/md_env.cu:1757-1778
        MDX_TRY(write_scalar(s, &DevScalars::gamma, s.Lx / s.Ly));
        MDX_TRY(write_scalar(s, &DevScalars::p_A,
                             std::max(s.p.barostat_mass, 1e-12) * s.p.barostat_area_rate_init));
        s.nph_init = true;
    }
    if (!s.K_valid) MDX_TRY(kinetic(s));
    if (!s.W_valid || !s.cells_valid || s.cells_py != 1) {
        MDX_TRY(ensure_cells_for(s, 1));
        MDX_TRY(launch_pair<FM_OBS, true, false>(s, s.grid, pair_args(s, 0.0)));
    }
    const PairArgs pa = pair_args(s, p_ext);
    k_box_nph<<<1, 64, 0, s.stream>>>(s.p.dt, p_ext, pa.W_baro, s.sc);
    MDX_HIP(hipGetLastError());
    DevScalars h;
    MDX_TRY(read_scalars(s, h));
    if (h.box_error) return fail(MDX_ERR_BOX, "step_single_NPH: invalid proposed area");
    s.Lx = h.Lx;
    s.Ly = h.Ly;
    const Grid g = grid_for(s, s.Lx, s.Ly, 1);
    MDX_TRY(ensure_cells(s, g));
    MDX_TRY(launch_drift<DK_NPH>(s, g, drift_args(s, 0)));
    MDX_TRY(launch_cells(s, g, 0, nullptr));
    MDX_TRY(launch_pair<FM_NPH, true, false>(s, g, pa));
    s.grid = g;
    s.cells_valid = true;
    s.cells_py = 1;
    s.K_valid = true;
    s.W_valid = true;
    s.t += s.p.dt;
    return MDX_OK;
}

int init_piston(mdx_sim &s) {   // src/md_env.cu:1780-1793
    if (!s.vp_init) {
        MDX_TRY(write_scalar(s, &DevScalars::v_p, 0.0));
        s.vp_init = true;
    }
    if (!s.piston_active) {
        s.piston_active = true;
        MDX_TRY(compute_forces_piston(s));
        DevScalars h;
        MDX_TRY(read_scalars(s, h));
        MDX_TRY(write_scalar(s, &DevScalars::last_P, (s.Lx > 0.0) ? h.F_p / s.Lx : 0.0));
    }
    return MDX_OK;
}

int step_piston(mdx_sim &s, double p_ext) {   // src/md_env.cu:1920-2023
    MDX_TRY(init_piston(s));
    const PairArgs pa = pair_args(s, p_ext);
    k_box_piston<<<1, 64, 0, s.stream>>>(s.p.dt, p_ext, pa.W_baro, s.sc);
    MDX_HIP(hipGetLastError());
    DevScalars h;
    MDX_TRY(read_scalars(s, h));
    if (h.box_error) return fail(MDX_ERR_BOX, "step_single_NPH_piston: invalid box");
    s.Ly = h.Ly;
    const Grid g = grid_for(s, s.Lx, s.Ly, 0);
    MDX_TRY(ensure_cells(s, g));
    MDX_TRY(launch_drift<DK_PISTON>(s, g, drift_args(s, 0)));
    MDX_TRY(launch_cells(s, g, 0, nullptr));
    MDX_TRY(launch_pair<FM_PISTON, false, true>(s, g, pa));
    s.grid = g;
    s.cells_valid = true;
    s.cells_py = 0;
    s.K_valid = false;
    s.W_valid = false;
    s.t += s.p.dt;
    return MDX_OK;
}

// U of the current positions (compute_U_energy_local), OBS pass.
int obs_pass(mdx_sim &s, int periodic_y) {
    MDX_TRY(ensure_cells_for(s, periodic_y));
    MDX_TRY(launch_pair<FM_OBS, true, false>(s, s.grid, pair_args(s, 0.0)));
    return MDX_OK;
}

int walls_pass(mdx_sim &s) {
    k_walls<<<blocks_for(s.n), kBlock, 0, s.stream>>>(static_cast<int>(s.n), s.x, s.Ly, s.p.k_piston, s.sc,
                                                       s.partials);
    MDX_HIP(hipGetLastError());
    return MDX_OK;
}

int validate_params(const mdx_params &p) {
    auto pos = [](double v) { return v > 0.0 && std::isfinite(v); };
    if (!pos(p.Lx) || !pos(p.Ly)) return fail(MDX_ERR_BOX, "box sides must be positive and finite");
    if (!pos(p.sigma_aa) || !pos(p.sigma_bb) || !pos(p.sigma_ab)) return fail(MDX_ERR_INVALID, "sigma must be > 0");
    if (!pos(p.mass_a) || !pos(p.mass_b)) return fail(MDX_ERR_INVALID, "masses must be > 0");
    if (!pos(p.cutoff)) return fail(MDX_ERR_INVALID, "cutoff must be > 0");
    if (!std::isfinite(p.dt)) return fail(MDX_ERR_INVALID, "dt must be finite");
    if (p.precision < MDX_PREC_COMP32 || p.precision > MDX_PREC_FP64)
        return fail(MDX_ERR_INVALID, "unknown precision");
    return MDX_OK;
}

int upload_frame(mdx_sim &s, const mdx_particle64 *frame) {
    const int64_t n = s.n;
    std::vector<double2> hx(n), hv(n);
    std::vector<uint8_t> hc(n);
    s.type.resize(n);
    s.theta.resize(n);
    for (int64_t i = 0; i < n; ++i) {
        hx[i] = make_double2(frame[i].pos[0], frame[i].pos[1]);
        hv[i] = make_double2(frame[i].vel[0], frame[i].vel[1]);
        const int t = frame[i].type;
        hc[i] = static_cast<uint8_t>(t == 0 ? 0 : (t == 1 ? 1 : 2));
        s.type[i] = t;
        s.theta[i] = frame[i].theta;
    }
    MDX_HIP(hipMemcpy(s.x, hx.data(), n * sizeof(double2), hipMemcpyHostToDevice));
    MDX_HIP(hipMemcpy(s.v, hv.data(), n * sizeof(double2), hipMemcpyHostToDevice));
    MDX_HIP(hipMemcpy(s.cls, hc.data(), n, hipMemcpyHostToDevice));
    s.cells_valid = false;
    s.K_valid = false;
    s.W_valid = false;
    return MDX_OK;
}

void free_sim(mdx_sim *s) {
    if (!s) return;
    void *ptrs[] = {s->x, s->v, s->a, s->cls, s->key, s->sorted, s->scell, s->img,
                    s->icls, s->partials, s->sc, s->count, s->cs, s->bsum};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (auto &e : s->force_ev) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    for (auto &e : s->step_ev) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

}  // namespace

extern "C" {

const char *mdx_last_error(void) { return g_err.c_str(); }

const char *mdx_version(void) { return "mdx 0.1 (gfx950)"; }

int mdx_device_count(int32_t *count) {
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) c = 0;
    if (count) *count = c;
    return MDX_OK;
}

int mdx_create(const mdx_params *params, const mdx_particle64 *frame, int64_t n, mdx_sim **out) {
    if (!params || !frame || !out) return fail(MDX_ERR_INVALID, "null argument");
    *out = nullptr;
    if (n <= 0) return fail(MDX_ERR_INVALID, "n must be positive");
    if (n > (int64_t{1} << 30)) return fail(MDX_ERR_CAPACITY, "n above 2^30");
    MDX_TRY(validate_params(*params));
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(MDX_ERR_NODEVICE, "no HIP device");
    mdx_sim *s = new mdx_sim();
    s->p = *params;
    if (!(s->p.cell_scale >= 1.0)) s->p.cell_scale = 1.0;
    s->n = n;
    s->prec = params->precision;
    s->Lx = s->Lx0 = params->Lx;
    s->Ly = s->Ly0 = params->Ly;
    make_pair_table(s->p, s->pt, s->rc_max);
    auto bail = [&](int rc) {
        free_sim(s);
        return rc;
    };
    if (params->device >= 0) {
        if (params->device >= ndev) return bail(fail(MDX_ERR_INVALID, "device ordinal out of range"));
        if (hipSetDevice(params->device) != hipSuccess) return bail(fail(MDX_ERR_HIP, "hipSetDevice failed"));
    }
    if (hipGetDevice(&s->device) != hipSuccess) return bail(fail(MDX_ERR_HIP, "hipGetDevice failed"));
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(MDX_ERR_HIP, "hipStreamCreate failed"));
    const int nb = blocks_for(n);
    bool ok = hipMalloc(&s->x, n * sizeof(double2)) == hipSuccess &&
              hipMalloc(&s->v, n * sizeof(double2)) == hipSuccess &&
              hipMalloc(&s->a, n * 16) == hipSuccess && hipMalloc(&s->cls, n) == hipSuccess &&
              hipMalloc(&s->key, n * 4) == hipSuccess && hipMalloc(&s->sorted, n * 4) == hipSuccess &&
              hipMalloc(&s->scell, n * 4) == hipSuccess && hipMalloc(&s->img, n * 16) == hipSuccess &&
              hipMalloc(&s->icls, n) == hipSuccess &&
              hipMalloc(&s->partials, 3 * static_cast<size_t>(nb) * sizeof(double)) == hipSuccess &&
              hipMalloc(&s->sc, sizeof(DevScalars)) == hipSuccess;
    if (!ok) return bail(fail(MDX_ERR_HIP, "device allocation failed"));
    DevScalars h{};
    h.Lx = s->Lx;
    h.Ly = s->Ly;
    h.gamma = s->Lx / s->Ly;
    h.Ly_floor = std::max(1e-12, s->p.barostat_height_min_ratio * s->Ly0);
    h.Ly_ceil = std::max(h.Ly_floor, s->p.barostat_height_max_ratio * s->Ly0);
    if (hipMemcpy(s->sc, &h, sizeof(h), hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(MDX_ERR_HIP, "scalar upload failed"));
    if (hipMemset(s->a, 0, n * 16) != hipSuccess) return bail(fail(MDX_ERR_HIP, "memset failed"));
    int rc = upload_frame(*s, frame);
    if (rc != MDX_OK) return bail(rc);
    rc = compute_forces(*s, 1);   // src/md_env.cu:66
    if (rc != MDX_OK) return bail(rc);
    if (hipStreamSynchronize(s->stream) != hipSuccess) return bail(fail(MDX_ERR_HIP, "sync failed"));
    *out = s;
    return MDX_OK;
}

int mdx_destroy(mdx_sim *sim) {
    if (!sim) return MDX_OK;
    (void)hipSetDevice(sim->device);
    (void)hipStreamSynchronize(sim->stream);
    free_sim(sim);
    return MDX_OK;
}

int mdx_set_state(mdx_sim *sim, const mdx_particle64 *frame, int64_t n) {
    if (!sim || !frame) return fail(MDX_ERR_INVALID, "null argument");
    if (n != sim->n)
        return fail(MDX_ERR_INVALID, "loaded particle count " + std::to_string(n) + " doesn't match n " +
                                         std::to_string(sim->n));
    MDX_HIP(hipSetDevice(sim->device));
    MDX_HIP(hipStreamSynchronize(sim->stream));
    MDX_TRY(upload_frame(*sim, frame));
    sim->piston_active = false;
    MDX_TRY(compute_forces(*sim, 1));
    MDX_HIP(hipStreamSynchronize(sim->stream));
    return MDX_OK;
}

int mdx_step(mdx_sim *sim, int32_t kind, int32_t nsteps, double p_ext, uint32_t flags) {
    if (!sim) return fail(MDX_ERR_INVALID, "null handle");
    if (nsteps < 0) return fail(MDX_ERR_INVALID, "nsteps < 0");
    MDX_HIP(hipSetDevice(sim->device));
    MDX_TRY(check_n(*sim));
    timing_begin(*sim, sim->step_ev);
    for (int k = 0; k < nsteps; ++k) {
        switch (kind) {
        case MDX_NVE: MDX_TRY(step_nve(*sim)); break;
        case MDX_NVT: MDX_TRY(step_nvt(*sim, flags)); break;
        case MDX_NPH: MDX_TRY(step_nph(*sim, p_ext)); break;
        case MDX_PISTON: MDX_TRY(step_piston(*sim, p_ext)); break;
        default: return fail(MDX_ERR_INVALID, "unknown ensemble");
        }
    }
    timing_end(*sim, sim->step_ev);
    if (sim->timing) sim->step_n += nsteps;
    return MDX_OK;
}

int mdx_observables(mdx_sim *sim, uint32_t mask, mdx_obs *out) {
    if (!sim || !out) return fail(MDX_ERR_INVALID, "null argument");
    MDX_HIP(hipSetDevice(sim->device));
    mdx_sim &s = *sim;
    const bool want_k = mask & (MDX_OBS_K | MDX_OBS_P);
    if (want_k && !s.K_valid) MDX_TRY(kinetic(s));
    double wallE = 0.0;
    if (s.piston_active) {
        if (mask & (MDX_OBS_U | MDX_OBS_VIRIAL)) MDX_TRY(obs_pass(s, 0));
        if (mask & (MDX_OBS_U | MDX_OBS_P)) MDX_TRY(walls_pass(s));
    } else if (mask & (MDX_OBS_U | MDX_OBS_P | MDX_OBS_VIRIAL)) {
        MDX_TRY(obs_pass(s, 1));
    }
    DevScalars h;
    MDX_TRY(read_scalars(s, h));
    std::memset(out, 0, sizeof(*out));
    out->K = h.K;
    if (s.piston_active) {
        wallE = h.wallE;
        if (mask & MDX_OBS_U) out->U = h.U + wallE;
        if (mask & MDX_OBS_VIRIAL) out->W = h.W;
        if (mask & MDX_OBS_P) {   // src/md_env.cu:2255-2261
            const double P = (s.Lx > 0.0) ? h.F_p / s.Lx : 0.0;
            MDX_TRY(write_scalar(s, &DevScalars::last_P, P));
            h.last_P = P;
            out->P = P;
        }
    } else {
        if (mask & MDX_OBS_U) out->U = h.U;
        if (mask & MDX_OBS_VIRIAL) out->W = h.W;
        if (mask & MDX_OBS_P) {   // src/md_env.cu:2205-2252
            const double area = s.Lx * s.Ly;
            const double P = (area > 0.0) ? (2.0 * h.K + h.W) / (2.0 * area) : 0.0;
            MDX_TRY(write_scalar(s, &DevScalars::last_P, P));
            h.last_P = P;
            out->P = P;
        }
    }
    out->xi = h.xi;
    out->p_A = h.p_A;
    out->area_rate = h.p_A / std::max(s.p.barostat_mass, 1e-12);
    out->v_p = h.v_p;
    out->F_p = h.F_p;
    out->last_P = h.last_P;
    out->Lx = s.Lx;
    out->Ly = s.Ly;
    out->t = s.t;
    out->piston_active = s.piston_active ? 1 : 0;
    return MDX_OK;
}

int mdx_partial_U_lambda(mdx_sim *sim, double eps_lambda, double *out) {
    if (!sim || !out) return fail(MDX_ERR_INVALID, "null argument");
    MDX_HIP(hipSetDevice(sim->device));
    MDX_TRY(obs_pass(*sim, sim->piston_active ? 0 : 1));
    DevScalars h;
    MDX_TRY(read_scalars(*sim, h));
    *out = -eps_lambda * h.lam;
    return MDX_OK;
}

int mdx_collect(mdx_sim *sim, mdx_particle64 *out, int64_t cap) {
    if (!sim || !out) return fail(MDX_ERR_INVALID, "null argument");
    if (cap < sim->n) return fail(MDX_ERR_CAPACITY, "output capacity below n");
    MDX_HIP(hipSetDevice(sim->device));
    const int64_t n = sim->n;
    std::vector<double2> hx(n), hv(n);
    std::vector<unsigned char> ha(n * 16);
    MDX_HIP(hipMemcpyAsync(hx.data(), sim->x, n * sizeof(double2), hipMemcpyDeviceToHost, sim->stream));
    MDX_HIP(hipMemcpyAsync(hv.data(), sim->v, n * sizeof(double2), hipMemcpyDeviceToHost, sim->stream));
    MDX_HIP(hipMemcpyAsync(ha.data(), sim->a, n * 16, hipMemcpyDeviceToHost, sim->stream));
    MDX_HIP(hipStreamSynchronize(sim->stream));
    const bool f32 = sim->prec == MDX_PREC_COMP32;
    for (int64_t i = 0; i < n; ++i) {
        mdx_particle64 &q = out[i];
        std::memset(&q, 0, sizeof(q));
        q.pos[0] = hx[i].x;
        q.pos[1] = hx[i].y;
        q.vel[0] = hv[i].x;
        q.vel[1] = hv[i].y;
        if (f32) {
            float fa[2];
            std::memcpy(fa, &ha[i * 8], 8);
            q.acc[0] = fa[0];
            q.acc[1] = fa[1];
        } else {
            std::memcpy(q.acc, &ha[i * 16], 16);
        }
        q.type = sim->type[i];
        q.theta = sim->theta[i];
    }
    return MDX_OK;
}

int mdx_compute_forces(mdx_sim *sim, int32_t periodic_y, double *acc_out) {
    if (!sim) return fail(MDX_ERR_INVALID, "null handle");
    MDX_HIP(hipSetDevice(sim->device));
    MDX_TRY(compute_forces(*sim, periodic_y ? 1 : 0));
    if (acc_out) {
        std::vector<mdx_particle64> tmp(sim->n);
        MDX_TRY(mdx_collect(sim, tmp.data(), sim->n));
        for (int64_t i = 0; i < sim->n; ++i) {
            acc_out[2 * i] = tmp[i].acc[0];
            acc_out[2 * i + 1] = tmp[i].acc[1];
        }
    }
    MDX_HIP(hipStreamSynchronize(sim->stream));
    return MDX_OK;
}

int mdx_fire(mdx_sim *sim, const mdx_fire_params *fp, double force_tol, int32_t max_steps,
             mdx_fire_result *result, int32_t *energy_steps, double *energy_trace, int32_t cap) {
    if (!sim || !fp || !result) return fail(MDX_ERR_INVALID, "null argument");
    if (max_steps <= 0) return fail(MDX_ERR_INVALID, "EnergyMinimizer: max_steps must be positive.");
    if (cap > 0 && (!energy_steps || !energy_trace)) return fail(MDX_ERR_INVALID, "null energy buffers");
    MDX_HIP(hipSetDevice(sim->device));
    mdx_sim &s = *sim;
    s.piston_active = false;
    // initialize_from_frame: wrap, v = a = 0 (src/md_energy_minimizer.cu:343-380)
    k_fire_init<<<blocks_for(s.n), kBlock, 0, s.stream>>>(static_cast<int>(s.n), s.x, s.v,
                                                           static_cast<double2 *>(s.a), s.Lx, s.Ly);
    MDX_HIP(hipGetLastError());
    // Overlapping starts produce accelerations beyond FP32 range: FIRE keeps
    // FP64 accelerations (FP32 pairs with FP64 accumulation, or FP64 pairs).
    if (s.prec == MDX_PREC_COMP32) s.prec = MDX_PREC_ACC64;
    DevScalars h;
    MDX_TRY(read_scalars(s, h));
    h.f_dt = fp->dt_init;
    h.f_alpha = fp->alpha0;
    h.f_maxF = 0.0;
    h.f_power = 0.0;
    h.f_streak = 0;
    h.f_done = 0;
    h.f_zero_v = 0;
    h.f_step = 0;
    h.f_conv_step = 0;
    MDX_HIP(hipMemcpyAsync(s.sc, &h, sizeof(h), hipMemcpyHostToDevice, s.stream));
    MDX_TRY(rebuild_cells(s, 1));
    PairArgs pa = pair_args(s, 0.0);
    pa.ftol = force_tol;
    pa.f_inc = fp->f_inc;
    pa.f_dec = fp->f_dec;
    pa.dt_max = fp->dt_max;
    pa.alpha0 = fp->alpha0;
    pa.N_min = fp->N_min;
    DriftArgs da{};
    da.max_step = fp->max_step;
    int ns = 0;
    int last_pushed = -1;
    auto push = [&](int step, double U) {
        if (ns < cap) {
            energy_steps[ns] = step;
            energy_trace[ns] = U;
        }
        ++ns;
        last_pushed = step;
    };
    auto potential = [&](double &U) -> int {
        MDX_TRY(obs_pass(s, 1));
        DevScalars q;
        MDX_TRY(read_scalars(s, q));
        U = q.U;
        return MDX_OK;
    };
    double U = 0.0;
    MDX_TRY(potential(U));   // src/md_energy_minimizer.cu:145-147
    push(0, U);
    const int interval = fp->energy_interval > 0 ? fp->energy_interval : 1000;
    const int *skip = &s.sc->f_done;
    int iters = 0;
    bool converged = false;
    const int kBatch = 128;
    while (iters < max_steps) {
        int end = std::min(max_steps, (iters / interval + 1) * interval);
        end = std::min(end, iters + kBatch);
        const Grid g = s.grid;
        for (int it = iters; it < end; ++it) {
            MDX_TRY(launch_pair<FM_FIRE, false, false>(s, g, pa));
            MDX_TRY(launch_drift<DK_FIRE>(s, g, da));
            MDX_TRY(launch_cells(s, g, 0, skip));
        }
        iters = end;
        MDX_TRY(read_scalars(s, h));
        if (h.f_done) {
            converged = true;
            break;
        }
        if (iters % interval == 0) {   // src/md_energy_minimizer.cu:209-215
            MDX_TRY(potential(U));
            push(iters, U);
        }
        if (fp->fp64_switch > 0.0 && s.prec != MDX_PREC_FP64 && h.f_maxF < fp->fp64_switch) {
            s.prec = MDX_PREC_FP64;   // the image changes representation; a stays FP64
            MDX_TRY(rebuild_cells(s, 1));
        }
    }
    int steps = max_steps;
    if (converged) {   // src/md_energy_minimizer.cu:170-178
        steps = h.f_conv_step;
        MDX_TRY(potential(U));
        push(steps + 1, U);
    } else if (last_pushed != max_steps) {   // :227-233
        MDX_TRY(potential(U));
        push(max_steps, U);
    }
    result->steps = steps;
    result->converged = converged ? 1 : 0;
    result->max_force = h.f_maxF;
    MDX_TRY(potential(U));   // :270
    result->potential_energy = U;
    result->n_energy = ns;
    s.K_valid = false;
    s.W_valid = false;
    return MDX_OK;
}

int mdx_get_n(const mdx_sim *sim, int64_t *n) {
    if (!sim || !n) return fail(MDX_ERR_INVALID, "null argument");
    *n = sim->n;
    return MDX_OK;
}

int mdx_set_precision(mdx_sim *sim, int32_t precision) {
    if (!sim) return fail(MDX_ERR_INVALID, "null handle");
    if (precision < MDX_PREC_COMP32 || precision > MDX_PREC_FP64) return fail(MDX_ERR_INVALID, "unknown precision");
    if (precision == sim->prec) return MDX_OK;
    MDX_HIP(hipSetDevice(sim->device));
    sim->prec = precision;
    // accelerations and the cell image change representation: recompute both
    MDX_TRY(compute_forces(*sim, sim->piston_active ? 0 : 1));
    MDX_HIP(hipStreamSynchronize(sim->stream));
    return MDX_OK;
}

int mdx_set_dt(mdx_sim *sim, double dt) {
    if (!sim) return fail(MDX_ERR_INVALID, "null handle");
    if (!std::isfinite(dt)) return fail(MDX_ERR_INVALID, "dt must be finite");
    sim->p.dt = dt;
    return MDX_OK;
}

int mdx_sync(mdx_sim *sim) {
    if (!sim) return fail(MDX_ERR_INVALID, "null handle");
    MDX_HIP(hipSetDevice(sim->device));
    MDX_HIP(hipStreamSynchronize(sim->stream));
    return MDX_OK;
}

int mdx_cell_list(mdx_sim *sim, int32_t periodic_y, int32_t *ncx, int32_t *ncy, uint32_t *keys,
                  uint32_t *cell_start, uint32_t *sorted, int64_t cap_n, int64_t cap_cells) {
    if (!sim) return fail(MDX_ERR_INVALID, "null handle");
    MDX_HIP(hipSetDevice(sim->device));
    MDX_TRY(rebuild_cells(*sim, periodic_y ? 1 : 0));
    const Grid &g = sim->grid;
    const int64_t ncell = static_cast<int64_t>(g.ncx) * g.ncy;
    if (ncx) *ncx = g.ncx;
    if (ncy) *ncy = g.ncy;
    if ((keys || sorted) && cap_n < sim->n) return fail(MDX_ERR_CAPACITY, "cap_n below n");
    if (cell_start && cap_cells < ncell + 1) return fail(MDX_ERR_CAPACITY, "cap_cells below ncell+1");
    if (keys) MDX_HIP(hipMemcpyAsync(keys, sim->key, sim->n * 4, hipMemcpyDeviceToHost, sim->stream));
    if (cell_start)
        MDX_HIP(hipMemcpyAsync(cell_start, sim->cs, (ncell + 1) * 4, hipMemcpyDeviceToHost, sim->stream));
    if (sorted) MDX_HIP(hipMemcpyAsync(sorted, sim->sorted, sim->n * 4, hipMemcpyDeviceToHost, sim->stream));
    MDX_HIP(hipStreamSynchronize(sim->stream));
    return MDX_OK;
}

int mdx_timing_enable(mdx_sim *sim, int32_t enable) {
    if (!sim) return fail(MDX_ERR_INVALID, "null handle");
    MDX_HIP(hipSetDevice(sim->device));
    MDX_HIP(hipStreamSynchronize(sim->stream));
    for (auto *vec : {&sim->force_ev, &sim->step_ev}) {
        for (auto &e : *vec) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        vec->clear();
    }
    sim->force_ms = sim->step_ms = 0.0;
    sim->force_n = sim->step_n = 0;
    sim->timing = enable != 0;
    return MDX_OK;
}

int mdx_timing_read(mdx_sim *sim, mdx_timing *out) {
    if (!sim || !out) return fail(MDX_ERR_INVALID, "null argument");
    MDX_HIP(hipSetDevice(sim->device));
    MDX_HIP(hipStreamSynchronize(sim->stream));
    auto drain = [](std::vector<std::pair<hipEvent_t, hipEvent_t>> &vec, double &ms, int64_t &cnt, bool count) {
        for (auto &e : vec) {
            float t = 0.f;
            if (hipEventElapsedTime(&t, e.first, e.second) == hipSuccess) {
                ms += t;
                if (count) ++cnt;
            }
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        vec.clear();
    };
    drain(sim->force_ev, sim->force_ms, sim->force_n, true);
    int64_t dummy = 0;
    drain(sim->step_ev, sim->step_ms, dummy, false);
    out->force_ms = sim->force_ms;
    out->force_launches = sim->force_n;
    out->step_ms = sim->step_ms;
    out->steps = sim->step_n;
    return MDX_OK;
}

}  // extern "C"
