// gfx950 kernels of the 2D binary Lennard-Jones step.
//
// Data layout (structure of arrays, original particle order unless noted):
//   x, v      double2[n]   FP64 master positions / velocities
//   a         float2[n] (compensated form) or double2[n] (FP64-accumulating forms)
//   cls       uint8[n]     type class: 0, 1, or 2 for any other type
//   key       uint32[n]    canonical cell index of each particle
//   count     uint32[ncell] per-cell occupancy (self-clearing: the scatter
//                          counts it back down to zero)
//   cs        uint32[ncell+1] cell_start (exclusive prefix of count)
//   sorted    uint32[n]    particle ids in cell order, ascending id per cell
//   img       float2[n] or double2[n], in cell order: cell-centre-relative FP32
//                          position, or the wrapped FP64 position
//   icls, scell           class and cell of each cell-ordered slot
//
// One step = drift kernel (kick/drift/wrap/cell key/count, plus the ensemble's
// reduction) -> two-pass scan -> scatter -> per-cell sort + image -> pair
// kernel (forces, second kick, the ensemble's reduction).  Global FP64 scalars
// (Nose-Hoover xi, NPH box and area momentum, piston, FIRE schedule) live in a
// device struct and are finalised by the last workgroup of a reduction.
#pragma once
#include "mdx_device.hpp"
#include "mdx_sim.hpp"

#include <math.h>

#include <type_traits>

namespace mdx {

// --------------------------------------------------------------------------
// drift kernels
// --------------------------------------------------------------------------

enum DriftKind { DK_KEYS = 0, DK_NVE = 1, DK_NVT = 2, DK_NPH = 3, DK_PISTON = 4, DK_FIRE = 5 };

template <typename AT>
__device__ inline double2 load_acc(const AT *a, int i);
template <>
__device__ inline double2 load_acc<float2>(const float2 *a, int i) {
    const float2 q = a[i];
    return make_double2(static_cast<double>(q.x), static_cast<double>(q.y));
}
template <>
__device__ inline double2 load_acc<double2>(const double2 *a, int i) {
    return a[i];
}

// Kick / drift / wrap of one particle by ensemble, then its cell key.
// Restates src/md_cuda_common.cu:381-496 (VV, NH, NPH scale+drift, piston)
// and src/md_energy_minimizer.cu:32-71 (FIRE zero/mix/drift).
template <int KIND, typename AT>
__global__ void __launch_bounds__(kBlock)
k_drift(int n, double2 *__restrict__ x, double2 *__restrict__ v, const AT *__restrict__ a,
        const uint8_t *__restrict__ cls, uint32_t *__restrict__ key, uint32_t *__restrict__ count,
        Grid g, DriftArgs da, PairTable pt, DevScalars *sc, double *partials) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if constexpr (KIND == DK_FIRE) {
        if (sc->f_done) return;
    }
    double red = 0.0;
    double xi1 = 0.0;
    if constexpr (KIND == DK_NVT) {
        xi1 = sc->xi + 0.5 * da.dt * (2.0 * sc->K - da.g_dof * da.T0) / da.Q;
    }
    if (i < n) {
        double2 p = x[i];
        const double dt = da.dt;
        if constexpr (KIND == DK_KEYS) {
            p.x = pbc_wrap(p.x, g.Lx);
            if (g.periodic_y) p.y = pbc_wrap(p.y, g.Ly);
        } else {
            double2 q = v[i];
            const double2 acc = load_acc<AT>(a, i);
            if constexpr (KIND == DK_NVE || KIND == DK_PISTON) {
                q.x += 0.5 * dt * acc.x;
                q.y += 0.5 * dt * acc.y;
                p.x += dt * q.x;
                p.y += dt * q.y;
                p.x = pbc_wrap(p.x, g.Lx);
                if (KIND == DK_NVE) p.y = pbc_wrap(p.y, g.Ly);
            } else if constexpr (KIND == DK_NVT) {
                double vx = q.x, vy = q.y;
                vx += 0.5 * dt * (acc.x - xi1 * vx);
                vy += 0.5 * dt * (acc.y - xi1 * vy);
                q.x = vx;
                q.y = vy;
                p.x += dt * vx;
                p.y += dt * vy;
                p.x = pbc_wrap(p.x, g.Lx);
                p.y = pbc_wrap(p.y, g.Ly);
                if (da.middle_wrap) {   // src/md_cuda_common.cu:26-65
                    double ym = fmod(p.x - da.slab_xmin, da.slab_period);
                    if (ym < 0.0) ym += da.slab_period;
                    if (ym > da.slab_width) ym = da.slab_period - ym;
                    p.x = da.slab_xmin + ym;
                }
            } else if constexpr (KIND == DK_NPH) {
                const double lin = sc->nph_lin, vsc = sc->nph_vsc;
                q.x += 0.5 * dt * acc.x;
                q.y += 0.5 * dt * acc.y;
                q.x *= vsc;
                q.y *= vsc;
                p.x = lin * p.x + dt * q.x;
                p.y = lin * p.y + dt * q.y;
                p.x = pbc_wrap(p.x, g.Lx);
                p.y = pbc_wrap(p.y, g.Ly);
            } else if constexpr (KIND == DK_FIRE) {
                if (sc->f_zero_v) {
                    q.x = 0.0;
                    q.y = 0.0;
                }
                const double alpha = sc->f_alpha;
                const double fdt = sc->f_dt;
                const double vmag = sqrt(q.x * q.x + q.y * q.y);
                const double fmag = sqrt(acc.x * acc.x + acc.y * acc.y);
                if (vmag > 0.0 && fmag > 0.0) {
                    const double s = vmag / fmag;
                    const double nvx = (1.0 - alpha) * q.x + alpha * s * acc.x;
                    const double nvy = (1.0 - alpha) * q.y + alpha * s * acc.y;
                    q.x = nvx;
                    q.y = nvy;
                }
                double ddx = fdt * q.x, ddy = fdt * q.y;
                if (da.max_step > 0.0) {
                    const double d = sqrt(ddx * ddx + ddy * ddy);
                    if (d > da.max_step) {
                        const double f = da.max_step / d;
                        ddx *= f;
                        ddy *= f;
                    }
                }
                p.x = pbc_wrap(p.x + ddx, g.Lx);
                p.y = pbc_wrap(p.y + ddy, g.Ly);
            }
            v[i] = q;
            x[i] = p;
            if constexpr (KIND == DK_NVT || KIND == DK_NPH) {
                const double m = pt.m[cls[i]];
                red = 0.5 * m * (q.x * q.x + q.y * q.y);
            }
        }
        const int cx = cell_coord(p.x, g.Lx, g.ncx);
        const int cy = cell_coord(p.y, g.Ly, g.ncy);
        const uint32_t c = static_cast<uint32_t>(cy) * static_cast<uint32_t>(g.ncx) + cx;
        key[i] = c;
        atomicAdd(&count[c], 1u);
    }
    if constexpr (KIND == DK_NVT || KIND == DK_NPH) {
        double vals[1] = {red};
        double out[1];
        if (grid_reduce<1, 0u>(vals, partials, &sc->ticket[0], out) && threadIdx.x == 0) {
            sc->K_half = out[0];
            if constexpr (KIND == DK_NVT)
                sc->xi_mid = xi1 + 0.5 * da.dt * (2.0 * out[0] - da.g_dof * da.T0) / da.Q;
        }
    }
}

// --------------------------------------------------------------------------
// cell list: scan, scatter, per-cell sort + image
// --------------------------------------------------------------------------

__global__ void __launch_bounds__(kBlock)
k_scan_a(const uint32_t *__restrict__ count, int ncell, uint32_t *__restrict__ bsum, const int *skip) {
    if (skip && *skip) return;
    __shared__ uint32_t sh[kWaves];
    const int base = blockIdx.x * kScanChunk + threadIdx.x * 4;
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j)
        if (base + j < ncell) s += count[base + j];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kWaves; ++w) t += sh[w];
        bsum[blockIdx.x] = t;
    }
}

__global__ void __launch_bounds__(kBlock)
k_scan_b(const uint32_t *__restrict__ count, int ncell, const uint32_t *__restrict__ bsum,
         uint32_t *__restrict__ cs, const int *skip) {
    if (skip && *skip) return;
    __shared__ uint32_t shw[kWaves];
    __shared__ uint32_t shp[kWaves];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    // prefix of the preceding chunks
    uint32_t pre = 0;
    for (int b = threadIdx.x; b < static_cast<int>(blockIdx.x); b += kBlock) pre += bsum[b];
    for (int o = 32; o > 0; o >>= 1) pre += __shfl_xor(pre, o, 64);
    if (lane == 0) shp[wid] = pre;
    const int base = blockIdx.x * kScanChunk + threadIdx.x * 4;
    uint32_t c[4];
    uint32_t tsum = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        c[j] = (base + j < ncell) ? count[base + j] : 0u;
        tsum += c[j];
    }
    // inclusive wave scan of thread sums
    uint32_t inc = tsum;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(inc, o, 64);
        if (lane >= o) inc += y;
    }
    if (lane == 63) shw[wid] = inc;
    __syncthreads();
    uint32_t chunk_pre = 0;
    for (int w = 0; w < kWaves; ++w) chunk_pre += shp[w];
    uint32_t wave_off = 0;
    for (int w = 0; w < wid; ++w) wave_off += shw[w];
    uint32_t run = chunk_pre + wave_off + inc - tsum;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (base + j < ncell) cs[base + j] = run;
        run += c[j];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == kBlock - 1) cs[ncell] = run;
}

__global__ void __launch_bounds__(kBlock)
k_scatter(int n, const uint32_t *__restrict__ key, uint32_t *__restrict__ count,
          const uint32_t *__restrict__ cs, uint32_t *__restrict__ sorted, const int *skip) {
    if (skip && *skip) return;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = key[i];
    const uint32_t slot = atomicSub(&count[c], 1u) - 1u;
    sorted[cs[c] + slot] = static_cast<uint32_t>(i);
}

// One thread per cell: restore ascending-id order inside the cell (the
// scatter's atomics leave it arbitrary), then write the cell-ordered image.
template <typename IT>
__global__ void __launch_bounds__(kBlock)
k_cell_image(int ncell, const uint32_t *__restrict__ cs, uint32_t *__restrict__ sorted,
             const double2 *__restrict__ x, const uint8_t *__restrict__ cls, Grid g, int wrap,
             IT *__restrict__ img, uint8_t *__restrict__ icls, uint32_t *__restrict__ scell,
             const int *skip) {
    if (skip && *skip) return;
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= ncell) return;
    const uint32_t s0 = cs[c], s1 = cs[c + 1];
    for (uint32_t k = s0 + 1; k < s1; ++k) {
        const uint32_t id = sorted[k];
        uint32_t j = k;
        while (j > s0) {
            const uint32_t prev = sorted[j - 1];
            if (prev <= id) break;
            sorted[j] = prev;
            --j;
        }
        sorted[j] = id;
    }
    const int cx = c % g.ncx, cy = c / g.ncx;
    const double ox = (static_cast<double>(cx) + 0.5) * g.w;
    const double oy = (static_cast<double>(cy) + 0.5) * g.h;
    for (uint32_t k = s0; k < s1; ++k) {
        const uint32_t id = sorted[k];
        double2 p = x[id];
        if (wrap) {
            p.x = pbc_wrap(p.x, g.Lx);
            if (g.periodic_y) p.y = pbc_wrap(p.y, g.Ly);
        }
        if constexpr (sizeof(IT) == sizeof(float2)) {
            img[k] = make_float2(static_cast<float>(p.x - ox), static_cast<float>(p.y - oy));
        } else {
            img[k] = p;
        }
        icls[k] = cls[id];
        scell[k] = static_cast<uint32_t>(c);
    }
}

// --------------------------------------------------------------------------
// pair kernel
// --------------------------------------------------------------------------

enum ForceMode { FM_PLAIN = 0, FM_OBS = 1, FM_NVE = 2, FM_NVT = 3, FM_NPH = 4, FM_PISTON = 5, FM_FIRE = 6 };
enum Acc { ACC_COMP32 = 0, ACC_FP64 = 1 };

template <typename IT>
struct ImgScalar;
template <>
struct ImgScalar<float2> { using T = float; };
template <>
struct ImgScalar<double2> { using T = double; };

// Pair accumulators of one particle.
template <int ACC>
struct ForceSum;
template <>
struct ForceSum<ACC_COMP32> {
    Comp32 fx{0.f, 0.f}, fy{0.f, 0.f};
    double gx = 0.0, gy = 0.0;   // FP64-guarded close pairs
    __device__ inline void add(float a, float b) { fx.add(a); fy.add(b); }
    __device__ inline void addd(double a, double b) { gx += a; gy += b; }
    __device__ inline double X() const { return fx.value() + gx; }
    __device__ inline double Y() const { return fy.value() + gy; }
};
template <>
struct ForceSum<ACC_FP64> {
    double fx = 0.0, fy = 0.0;
    __device__ inline void add(float a, float b) { fx += static_cast<double>(a); fy += static_cast<double>(b); }
    __device__ inline void addd(double a, double b) { fx += a; fy += b; }
    __device__ inline double X() const { return fx; }
    __device__ inline double Y() const { return fy; }
};

struct ObsSum {
    double u = 0.0, w = 0.0, lam = 0.0;
};

// FP64 evaluation of one pair from its displacement; the operation order is
// the oracle's (and the reference's, src/md_cuda_common.cu:237-254) so the
// FP64 form reproduces the oracle bit for bit.
template <bool OBS>
__device__ inline bool pair_fp64(double dx, double dy, int pidx, const PairTable &pt, double &fx,
                                 double &fy, ObsSum &ob) {
    const double r2 = dx * dx + dy * dy;
    if (!(r2 > 0.0 && r2 < pt.rc2[pidx])) return false;
    const double r2i = 1.0 / r2;
    const double sr2 = pt.sig2[pidx] * r2i;
    const double sr6 = sr2 * sr2 * sr2;
    const double sr12 = sr6 * sr6;
    const double tmp = pt.eps24[pidx] * (2.0 * sr12 - sr6) * r2i;
    fx = tmp * dx;
    fy = tmp * dy;
    if constexpr (OBS) {
        ob.u += 0.5 * (pt.eps4[pidx] * (sr12 - sr6) - pt.vc[pidx]);
        ob.w += 0.5 * (dx * fx + dy * fy);
        ob.lam += 0.5 * (fx * dx - fy * dy);
    }
    return true;
}

// Selects entry 0 (same class), 1 (other class) of a per-particle pair set.
struct PairSel {
    float rc2[2], sig2[2], eps24[2], eps4[2];
};


// Last-workgroup finalisation of the pair kernel's reduction out[0..2].
template <int MODE, bool OBS, bool WALLS>
__device__ inline void finalize_pair(const double (&out)[3], const PairArgs &pa, const Grid &g,
                                     DevScalars *sc) {
    if constexpr (MODE == FM_OBS) {
        sc->U = out[0];
        sc->W = out[1];
        sc->lam = out[2];
    } else if constexpr (MODE == FM_PLAIN) {
        if constexpr (OBS) {
            sc->U = out[0];
            sc->W = out[1];
            sc->lam = out[2];
        }
        if constexpr (WALLS) {   // cal_forces_piston, src/md_env.cu:1570-1573
            sc->F_p = out[1];
            sc->wallE = out[2];
        }
    } else if constexpr (MODE == FM_NVT) {
        sc->K = out[0];
        sc->xi = sc->xi_mid;
    } else if constexpr (MODE == FM_NPH) {   // src/md_env.cu:1896-1917
        const double W = out[1];
        const double area = g.Lx * g.Ly;
        const double P_new = (area > 0.0) ? (2.0 * sc->K_half + W) / (2.0 * area) : 0.0;
        double pA = sc->p_A + 0.5 * pa.dt * (P_new - pa.p_ext);
        if (sc->clamped) {
            if (g.Ly <= sc->Ly_floor && pA < 0.0) pA = 0.0;
            if (g.Ly >= sc->Ly_ceil && pA > 0.0) pA = 0.0;
        }
        sc->p_A = pA;
        sc->last_P = P_new;
        sc->K = out[0];
        sc->W = W;
    } else if constexpr (MODE == FM_PISTON) {   // src/md_env.cu:2004-2022
        sc->F_p = out[1];
        sc->wallE = out[2];
        const double acc_new = (sc->F_p - pa.p_ext * g.Lx) / pa.W_baro;
        double vp = sc->v_p + 0.5 * pa.dt * acc_new;
        if (sc->clamped) {
            if (g.Ly <= sc->Ly_floor && vp < 0.0) vp = 0.0;
            if (g.Ly >= sc->Ly_ceil && vp > 0.0) vp = 0.0;
        }
        sc->v_p = vp;
        sc->last_P = sc->F_p / g.Lx;
    } else if constexpr (MODE == FM_FIRE) {   // src/md_energy_minimizer.cu:170-203
        const double maxF = out[0], power = out[1];
        sc->f_maxF = maxF;
        sc->f_power = power;
        const int step = sc->f_step;
        if (maxF < pa.ftol) {
            sc->f_done = 1;
            sc->f_conv_step = step;
        } else if (power > 0.0) {
            sc->f_streak += 1;
            if (sc->f_streak > pa.N_min) {
                const double inc = sc->f_dt * pa.f_inc;
                sc->f_dt = (pa.dt_max < inc) ? pa.dt_max : inc;
                const double al = sc->f_alpha * 0.99;
                sc->f_alpha = (0.0 < al) ? al : 0.0;
            }
            sc->f_zero_v = 0;
        } else {
            sc->f_dt *= pa.f_dec;
            sc->f_streak = 0;
            sc->f_alpha = pa.alpha0;
            sc->f_zero_v = 1;
        }
        sc->f_step = step + 1;
    }
    (void)pa;
    (void)g;
}

template <typename IT, int ACC, int MODE, bool OBS, bool WALLS>
__global__ void __launch_bounds__(kBlock)
k_pair(int n, const uint32_t *__restrict__ cs, const IT *__restrict__ img,
       const uint8_t *__restrict__ icls, const uint32_t *__restrict__ scell,
       const uint32_t *__restrict__ sorted, Grid g, PairTable pt, PairArgs pa,
       typename std::conditional<ACC == ACC_COMP32, float2, double2>::type *__restrict__ a,
       double2 *__restrict__ v, const double2 *__restrict__ x, DevScalars *sc, double *partials) {
    using T = typename ImgScalar<IT>::T;
    using AT = typename std::conditional<ACC == ACC_COMP32, float2, double2>::type;
    constexpr bool kFp64 = sizeof(T) == sizeof(double);
    constexpr bool kReduce = OBS || WALLS || MODE == FM_NVT || MODE == FM_NPH || MODE == FM_FIRE;
    if constexpr (MODE == FM_FIRE) {
        if (sc->f_done) return;
    }
    const int s = blockIdx.x * kBlock + threadIdx.x;
    double r0 = 0.0, r1 = 0.0, r2v = 0.0;
    double xi_mid = 0.0;
    if constexpr (MODE == FM_NVT) xi_mid = sc->xi_mid;
    double fire_dt = 0.0;
    if constexpr (MODE == FM_FIRE) fire_dt = sc->f_dt;
    if (s < n) {
        const uint32_t c = scell[s];
        const int cx = static_cast<int>(c % static_cast<uint32_t>(g.ncx));
        const int cy = static_cast<int>(c / static_cast<uint32_t>(g.ncx));
        const IT pi = img[s];
        const int ci = icls[s];
        ForceSum<kFp64 ? ACC_FP64 : ACC> F;
        ObsSum ob;
        PairSel ps;
        if constexpr (!kFp64) {
            const int same = ci, other = 2;
            ps.rc2[0] = pt.rc2f[same];   ps.rc2[1] = pt.rc2f[other];
            ps.sig2[0] = pt.sig2f[same]; ps.sig2[1] = pt.sig2f[other];
            ps.eps24[0] = pt.eps24f[same]; ps.eps24[1] = pt.eps24f[other];
            ps.eps4[0] = pt.eps4f[same]; ps.eps4[1] = pt.eps4f[other];
        }
        for (int oy = -1; oy <= 1; ++oy) {
            int yy = cy + oy;
            T sy = 0;
            if (g.periodic_y) {
                if (yy < 0) { yy += g.ncy; sy = -static_cast<T>(g.Ly); }
                else if (yy >= g.ncy) { yy -= g.ncy; sy = static_cast<T>(g.Ly); }
            } else if (yy < 0 || yy >= g.ncy) {
                continue;
            }
            const T shy = kFp64 ? sy : static_cast<T>(oy) * static_cast<T>(g.hf);
            const int rowbase = yy * g.ncx;
            for (int ox = -1; ox <= 1; ++ox) {
                int xx = cx + ox;
                T sx = 0;
                if (xx < 0) { xx += g.ncx; sx = -static_cast<T>(g.Lx); }
                else if (xx >= g.ncx) { xx -= g.ncx; sx = static_cast<T>(g.Lx); }
                const T shx = kFp64 ? sx : static_cast<T>(ox) * static_cast<T>(g.wf);
                const uint32_t k0 = cs[rowbase + xx], k1 = cs[rowbase + xx + 1];
                for (uint32_t k = k0; k < k1; ++k) {
                    if (k == static_cast<uint32_t>(s)) continue;
                    const IT pj = img[k];
                    const int cj = icls[k];
                    if constexpr (kFp64) {
                        double dx = (pi.x - pj.x) - shx;
                        double dy = (pi.y - pj.y) - shy;
                        if (g.small_x) {
                            const double d = pi.x - pj.x;
                            dx = d - g.Lx * round(d / g.Lx);
                        }
                        if (g.periodic_y && g.small_y) {
                            const double d = pi.y - pj.y;
                            dy = d - g.Ly * round(d / g.Ly);
                        }
                        double fx, fy;
                        if (pair_fp64<OBS>(dx, dy, pair_index(ci, cj), pt, fx, fy, ob)) F.addd(fx, fy);
                    } else {
                        float dx = (pi.x - pj.x) - shx;
                        float dy = (pi.y - pj.y) - shy;
                        if (g.small_x) dx -= g.Lxf * rintf(dx / g.Lxf);
                        if (g.periodic_y && g.small_y) dy -= g.Lyf * rintf(dy / g.Lyf);
                        const float rr = dx * dx + dy * dy;
                        const int o = (cj == ci) ? 0 : 1;
                        const float rc2 = o ? ps.rc2[1] : ps.rc2[0];
                        if (rr < rc2 && rr > 0.0f) {
                            if (rr < pt.r2_guard) {
                                double fx, fy;
                                pair_fp64<OBS>(static_cast<double>(dx), static_cast<double>(dy),
                                               pair_index(ci, cj), pt, fx, fy, ob);
                                F.addd(fx, fy);
                            } else {
                                const float sig2 = o ? ps.sig2[1] : ps.sig2[0];
                                const float e24 = o ? ps.eps24[1] : ps.eps24[0];
                                const float r2i = 1.0f / rr;
                                const float sr2 = sig2 * r2i;
                                const float sr6 = sr2 * sr2 * sr2;
                                const float sr12 = sr6 * sr6;
                                const float tmp = e24 * (2.0f * sr12 - sr6) * r2i;
                                const float fx = tmp * dx, fy = tmp * dy;
                                F.add(fx, fy);
                                if constexpr (OBS) {
                                    const float e4 = o ? ps.eps4[1] : ps.eps4[0];
                                    const float u = e4 * (sr12 - sr6);
                                    const double ddx = dx, ddy = dy, dfx = fx, dfy = fy;
                                    ob.u += 0.5 * (static_cast<double>(u) - pt.vc[pair_index(ci, cj)]);
                                    ob.w += 0.5 * (ddx * dfx + ddy * dfy);
                                    ob.lam += 0.5 * (dfx * ddx - dfy * ddy);
                                }
                            }
                        }
                    }
                }
            }
        }
        const uint32_t id = sorted[s];
        double axd, ayd;
        if constexpr (kFp64 || ACC == ACC_FP64) {
            axd = F.X() / pt.m[ci];
            ayd = F.Y() / pt.m[ci];
        } else {
            axd = F.X() * pt.inv_m[ci];
            ayd = F.Y() * pt.inv_m[ci];
        }
        if constexpr (WALLS) {   // src/md_cuda_common.cu:498-552
            const double y = x[id].y;
            double fw = 0.0, up = 0.0, e = 0.0;
            if (pa.k_piston > 0.0) {
                if (y > g.Ly) {
                    const double ov = y - g.Ly;
                    fw -= pa.k_piston * ov;
                    up += pa.k_piston * ov;
                    e += 0.5 * pa.k_piston * ov * ov;
                }
                if (y < 0.0) {
                    const double ov = -y;
                    fw += pa.k_piston * ov;
                    e += 0.5 * pa.k_piston * ov * ov;
                }
            }
            if (fw != 0.0) ayd += fw / pt.m[ci];
            r1 = up;
            r2v = e;
        }
        if constexpr (MODE != FM_OBS) {
            AT out;
            out.x = static_cast<decltype(out.x)>(axd);
            out.y = static_cast<decltype(out.y)>(ayd);
            a[id] = out;
            // the kicks use the stored (possibly FP32-rounded) acceleration
            axd = static_cast<double>(out.x);
            ayd = static_cast<double>(out.y);
        }
        if constexpr (MODE == FM_NVE || MODE == FM_NPH || MODE == FM_PISTON) {
            double2 q = v[id];
            q.x += 0.5 * pa.dt * axd;
            q.y += 0.5 * pa.dt * ayd;
            v[id] = q;
            if constexpr (MODE == FM_NPH) r0 = 0.5 * pt.m[ci] * (q.x * q.x + q.y * q.y);
        } else if constexpr (MODE == FM_NVT) {
            double2 q = v[id];
            double vx = q.x, vy = q.y;
            vx += 0.5 * pa.dt * (axd - xi_mid * vx);
            vy += 0.5 * pa.dt * (ayd - xi_mid * vy);
            q.x = vx;
            q.y = vy;
            v[id] = q;
            r0 = 0.5 * pt.m[ci] * (vx * vx + vy * vy);
        } else if constexpr (MODE == FM_FIRE) {
            double2 q = v[id];
            q.x += fire_dt * axd;
            q.y += fire_dt * ayd;
            v[id] = q;
            r0 = sqrt(axd * axd + ayd * ayd);
            r1 = q.x * axd + q.y * ayd;
        }
        if constexpr (MODE == FM_NPH) {
            r1 = ob.w;
        } else if constexpr (OBS) {
            r0 = ob.u;
            r1 = ob.w;
            r2v = ob.lam;
        }
        (void)x;
    }
    if constexpr (kReduce) {
        double vals[3] = {r0, r1, r2v};
        double out[3];
        constexpr unsigned mm = (MODE == FM_FIRE) ? 1u : 0u;
        if (grid_reduce<3, mm>(vals, partials, &sc->ticket[1], out) && threadIdx.x == 0)
            finalize_pair<MODE, OBS, WALLS>(out, pa, g, sc);
    }
}

// --------------------------------------------------------------------------
// small kernels: kinetic energy, walls, box updates, FIRE set-up
// --------------------------------------------------------------------------

// K = sum 1/2 m v^2 (src/md_cuda_common.cu:554-589).
__global__ void __launch_bounds__(kBlock)
k_kinetic(int n, const double2 *__restrict__ v, const uint8_t *__restrict__ cls, PairTable pt,
          DevScalars *sc, double *partials) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double k = 0.0;
    if (i < n) {
        const double2 q = v[i];
        k = 0.5 * pt.m[cls[i]] * (q.x * q.x + q.y * q.y);
    }
    double vals[1] = {k};
    double out[1];
    if (grid_reduce<1, 0u>(vals, partials, &sc->ticket[2], out) && threadIdx.x == 0) sc->K = out[0];
}

// Piston wall observables without touching accelerations
// (src/md_env.cu:2148-2203 with update_particle_acc = false).
__global__ void __launch_bounds__(kBlock)
k_walls(int n, const double2 *__restrict__ x, double Ly, double k_piston, DevScalars *sc,
        double *partials) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    double up = 0.0, e = 0.0;
    if (i < n && k_piston > 0.0) {
        const double y = x[i].y;
        if (y > Ly) {
            const double ov = y - Ly;
            up += k_piston * ov;
            e += 0.5 * k_piston * ov * ov;
        }
        if (y < 0.0) {
            const double ov = -y;
            e += 0.5 * k_piston * ov * ov;
        }
    }
    double vals[2] = {up, e};
    double out[2];
    if (grid_reduce<2, 0u>(vals, partials, &sc->ticket[3], out) && threadIdx.x == 0) {
        sc->F_p = out[0];
        sc->wallE = out[1];
    }
}

// NPH box update of src/md_env.cu:1824-1877, one thread.
__global__ void k_box_nph(double dt, double p_ext, double W_baro, DevScalars *sc) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double Lx_old = sc->Lx, Ly_old = sc->Ly;
    const double area_old = Lx_old * Ly_old;
    sc->box_error = 0;
    if (!(area_old > 0.0)) {
        sc->box_error = 1;
        return;
    }
    const double P_old = (2.0 * sc->K + sc->W) / (2.0 * area_old);
    sc->P_old = P_old;
    double pA = sc->p_A + 0.5 * dt * (P_old - p_ext);
    double area_new = area_old + dt * (pA / W_baro);
    if (!(area_new > 0.0) || !isfinite(area_new)) {
        sc->box_error = 2;
        sc->p_A = pA;
        return;
    }
    double gam = sc->gamma;
    if (!(gam > 0.0) || !isfinite(gam)) gam = Lx_old / Ly_old;
    double Ly_new = sqrt(area_new / gam);
    double Lx_new = gam * Ly_new;
    int clamped = 0;
    if (Ly_new < sc->Ly_floor) {
        Ly_new = sc->Ly_floor;
        Lx_new = gam * Ly_new;
        area_new = Lx_new * Ly_new;
        clamped = 1;
    } else if (Ly_new > sc->Ly_ceil) {
        Ly_new = sc->Ly_ceil;
        Lx_new = gam * Ly_new;
        area_new = Lx_new * Ly_new;
        clamped = 1;
    }
    const double lin = sqrt(area_new / area_old);
    sc->nph_lin = lin;
    sc->nph_vsc = 1.0 / lin;
    sc->gamma = gam;
    sc->p_A = pA;
    sc->clamped = clamped;
    sc->Lx = Lx_new;
    sc->Ly = Ly_new;
}

// Piston update of src/md_env.cu:1950-1980, one thread.
__global__ void k_box_piston(double dt, double p_ext, double W_baro, DevScalars *sc) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double Lx = sc->Lx, Ly_old = sc->Ly;
    const double acc_old = (sc->F_p - p_ext * Lx) / W_baro;
    double vp = sc->v_p + 0.5 * dt * acc_old;
    double Ly_new = Ly_old + dt * vp;
    int clamped = 0;
    if (Ly_new < sc->Ly_floor) {
        Ly_new = sc->Ly_floor;
        clamped = 1;
        if (vp < 0.0) vp = 0.0;
    } else if (Ly_new > sc->Ly_ceil) {
        Ly_new = sc->Ly_ceil;
        clamped = 1;
        if (vp > 0.0) vp = 0.0;
    }
    sc->v_p = vp;
    sc->clamped = clamped;
    sc->Ly = Ly_new;
    sc->box_error = (Ly_new > 0.0 && isfinite(Ly_new)) ? 0 : 3;
}

// EnergyMinimizer::initialize_from_frame (src/md_energy_minimizer.cu:343-380):
// wrap positions, zero velocities and accelerations (a buffer viewed as 16 B
// per particle whatever its precision).
__global__ void __launch_bounds__(kBlock)
k_fire_init(int n, double2 *__restrict__ x, double2 *__restrict__ v, double2 *__restrict__ a16,
            double Lx, double Ly) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    double2 p = x[i];
    p.x = pbc_wrap(p.x, Lx);
    p.y = pbc_wrap(p.y, Ly);
    x[i] = p;
    v[i] = make_double2(0.0, 0.0);
    a16[i] = make_double2(0.0, 0.0);
}

}  // namespace mdx
