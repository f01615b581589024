// gfx950 kernels of the 2D binary Lennard-Jones step, batched over slots.
//
// Data layout (structure of arrays over the concatenated particles of all
// slots; slot r owns [off_r, off_r + n_cap_r)):
//   x, v      double2       FP64 master positions / velocities
//   a         float2 (compensated FP32 form) or double2 (FP64-accumulating forms)
//   cls       uint8         type class: 0, 1, or 2 for any other type
//   key       uint32        global cell index of each particle
//   count     uint32[C]     per-cell stayers of the list being built
//   arrivals  uint32[C]     per-cell movers arriving this step (zero at rest)
//   cs        uint32[C+1]   cell start: exclusive prefix of count over all cells
//   sorted    uint32        particle ids in cell order, ascending id per cell
//   img       int2/double2, in cell order: cell-centre-relative position in
//                           fixed point (FP32 forms, Grid), or the wrapped
//                           FP64 position (FP64 form)
//   icls, scell             class and slot-local cell of each sorted slot
//
// One step = k_box (one workgroup per slot: the previous pair pass's
// reduction, then box, barostat, piston, grid, FIRE bookkeeping) -> k_drift
// (kick/drift/wrap, cell key, the ensemble's reduction partials, movers) ->
// k_stay_arrive (stayers per cell, movers' arrival index) ->
// k_scan_a/k_scan_b -> k_fill -> k_cell_sort (id order of the cells that
// received movers) -> k_slot_image (the drift reduction; per-slot image,
// coalesced) -> k_list_build (neighbour lists, list mode) -> k_pair (forces,
// second kick, the ensemble's reduction partials).  A lean step of a rebuild
// window is k_lean (the previous pair pass's reduction, redone by each of its
// workgroups, then kick/drift/image) -> k_pair.  A run of steps
// ends with k_finalize, which combines the last pair pass's partials.  The
// cell list is updated from the previous one: only particles whose cell
// changed ("movers") use atomics.  Global FP64 scalars are combined from the
// workgroup partials by one workgroup per slot, in a fixed order.
#pragma once
#include <math.h>

#include <type_traits>

#include "mdx_device.hpp"
#include "mdx_sim.hpp"

namespace mdx {

template <typename AT>
__device__ inline double2 load_acc(const void *a, int i) {
    if constexpr (std::is_same_v<AT, float2>) {
        const float2 q = static_cast<const float2 *>(a)[i];
        return make_double2(static_cast<double>(q.x), static_cast<double>(q.y));
    } else {
        return static_cast<const double2 *>(a)[i];
    }
}

// --------------------------------------------------------------------------
// per-slot scalar kernel
// --------------------------------------------------------------------------

__device__ inline void finalize_pair(int pmode, const double (&out)[3], const SysParams &sp, const Grid &g,
                                     DevScalars &sc);

// Combines the partials of the slot's last pair pass (one whole workgroup)
// and applies its mode's scalar update (finalize_pair).
__device__ inline void finalize_pair_pending(const BatchPtrs &P, const SysParams &sp, DevScalars &sc) {
    const int pmode = sp.pmode;
    double out[3];
    slot_reduce<3>(pmode == PM_FIRE ? 1u : 0u, P.partials, P.TB, sp.boff, act_blocks(sp), out);
    if (threadIdx.x == 0) {
        finalize_pair(pmode, out, sp, sc.g, sc);
        sc.pair_pending = 0;
    }
}

// One workgroup per slot, first kernel of a step.  It first combines the
// previous pass's pair reduction, then thread 0 updates the slot's box.
// NPH: area barostat of src/md_env.cu:1824-1877.  Piston:
// src/md_env.cu:1950-1980.  FIRE: decides whether the pending drift and the
// next force pass run.  Every slot: the cell grid of its (possibly new) box.
enum StepMode : int32_t {
    STEP_NORMAL = 0,    // list slots rebuild when a particle moved beyond skin / 2
    STEP_REBUILD = 1,   // list slots rebuild their cell and neighbour lists
    STEP_LEAN = 2,      // a rebuild window's lean step: k_lean and k_pair only
    STEP_HEAD = 3,      // first step of a window: STEP_NORMAL, with lists that
                        // must also hold through the window's lean steps
};

__device__ inline double nh_half_update(double xi, double K, const SysParams &sp);

__global__ void __launch_bounds__(kBlock) k_box(BatchPtrs P, int32_t mode) {
    const int r = blockIdx.x;
    const SysParams &sp = P.sp[r];
    DevScalars &sc = P.sc[r];
    // the padding cell anchors the slot's cell starts at its first particle
    if (threadIdx.x == 0) P.count[sp.coff + sp.ccap] = static_cast<uint32_t>(sp.n > 0 ? sp.n_cap - sp.n : sp.n_cap);
    if (sp.n <= 0) {   // unloaded: the cell-list kernels of one step clear every cell of the slot
        if (threadIdx.x == 0) {
            sc.rebuild = sc.cleared ? 0 : 1;   // after that the slot takes no part in them
            sc.cleared = 1;
            sc.list_valid = 0;
            sc.list_built = 0;
            sc.list_on = 0;
            sc.fallback[0] = sc.fallback[1] = 0;
        }
        return;
    }
    if (sc.pair_pending && sp.nblk > 0) finalize_pair_pending(P, sp, sc);
    if (threadIdx.x != 0) return;
    if (sp.kind == SK_NONE && sc.list_valid) {
        // an idle slot keeps its cell list (and neighbour lists) as they are:
        // its cells' counts hold their occupancy, so the scan keeps its cell
        // starts, and its next step continues from the same lists
        sc.rebuild = 0;
        sc.fallback[0] = sc.fallback[1] = 0;
        return;
    }
    double Lx = sc.g.Lx, Ly = sc.g.Ly;
    if (sp.kind == SK_NPH) {
        const double area_old = Lx * Ly;
        sc.box_error = 0;
        if (!(area_old > 0.0)) {
            sc.box_error = 1;
            sc.nph_lin = 1.0;
            sc.nph_vsc = 1.0;
        } else {
            const double P_old = (2.0 * sc.K + sc.W) / (2.0 * area_old);
            sc.P_old = P_old;
            const double pA = sc.p_A + 0.5 * sp.dt * (P_old - sp.p_ext);
            double area_new = area_old + sp.dt * (pA / sp.W_baro);
            sc.p_A = pA;
            if (!(area_new > 0.0) || !isfinite(area_new)) {
                sc.box_error = 2;   // the reference aborts here
                sc.nph_lin = 1.0;
                sc.nph_vsc = 1.0;
            } else {
                double gam = sc.gamma;
                if (!(gam > 0.0) || !isfinite(gam)) gam = Lx / Ly;
                double Ly_new = sqrt(area_new / gam);
                double Lx_new = gam * Ly_new;
                int clamped = 0;
                if (Ly_new < sc.Ly_floor) {
                    Ly_new = sc.Ly_floor;
                    Lx_new = gam * Ly_new;
                    area_new = Lx_new * Ly_new;
                    clamped = 1;
                } else if (Ly_new > sc.Ly_ceil) {
                    Ly_new = sc.Ly_ceil;
                    Lx_new = gam * Ly_new;
                    area_new = Lx_new * Ly_new;
                    clamped = 1;
                }
                const double lin = sqrt(area_new / area_old);
                sc.nph_lin = lin;
                sc.nph_vsc = 1.0 / lin;
                sc.lin_cum *= lin;
                sc.gamma = gam;
                sc.clamped = clamped;
                Lx = Lx_new;
                Ly = Ly_new;
            }
        }
    } else if (sp.kind == SK_PISTON) {
        const double acc_old = (sc.F_p - sp.p_ext * Lx) / sp.W_baro;
        double vp = sc.v_p + 0.5 * sp.dt * acc_old;
        double Ly_new = Ly + sp.dt * vp;
        int clamped = 0;
        if (Ly_new < sc.Ly_floor) {
            Ly_new = sc.Ly_floor;
            clamped = 1;
            if (vp < 0.0) vp = 0.0;
        } else if (Ly_new > sc.Ly_ceil) {
            Ly_new = sc.Ly_ceil;
            clamped = 1;
            if (vp > 0.0) vp = 0.0;
        }
        sc.v_p = vp;
        sc.clamped = clamped;
        Ly = Ly_new;
    } else if (sp.kind == SK_FIRE) {
        const int live = !sc.f_done;
        sc.f_do_drift = (live && sc.f_drifted < sc.f_step) ? 1 : 0;
        if (sc.f_do_drift) sc.f_drifted = sc.f_step;
        sc.f_do_pair = (live && sc.f_step < sp.f_max_steps) ? 1 : 0;
    }
    sc.g_old = sc.g;
    Grid g = make_grid(Lx, Ly, sp.cell_side, sp.rc_max, sp.periodic_y);
    double side = sp.cell_side;
    for (int it = 0; it < 64 && g.ncx * g.ncy > sp.cbound; ++it) {   // never for boxes within the load-time bound
        side *= 1.25;
        g = make_grid(Lx, Ly, side, sp.rc_max, sp.periodic_y);
    }
    sc.g = g;
    // the neighbour list needs cells of at least r_list in a periodic box;
    // outside list mode every step rebuilds the cell list
    const bool on = sp.list_mode && !sc.list_overflow && g.periodic_y && g.w >= sp.r_list && g.h >= sp.r_list;
    sc.list_on = on ? 1 : 0;
    // a barostat step that changes the grid's dimensions renumbers the cells
    // the lists refer to
    const bool regrid = g.ncx != sc.g_old.ncx || g.ncy != sc.g_old.ncy;
    sc.rebuild = (!on || !sc.list_built || mode == STEP_REBUILD || regrid) ? 1 : 0;
    sc.fallback[0] = sc.fallback[1] = 0;   // (lean steps: k_lean of the step before clears its own)
}

// Half update of the Nose-Hoover variable, src/md_env.cu:1668-1673:
// xi += 0.5 dt (2K - g T0) / Q.
__device__ inline double nh_half_update(double xi, double K, const SysParams &sp) {
    return xi + 0.5 * sp.dt * (2.0 * K - sp.g_dof * sp.T0) / sp.Q;
}

// Appends the calling lanes with mover == true to the batch's mover list: one
// atomic per wave that has movers.
__device__ inline void append_movers(const BatchPtrs &P, bool mover, uint32_t pid, int gb) {
    // Workgroup-local list: movers[gb * kBlock + j], j < n_movers[gb], in
    // lane order.  No global atomics (a single global counter serialises
    // across the XCDs' L2s).
    __shared__ uint32_t wcount[kWaves];
    const unsigned long long mask = __ballot(mover);
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    if (lane == 0) wcount[wid] = static_cast<uint32_t>(__popcll(mask));
    __syncthreads();
    uint32_t base = 0, total = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        base += w < wid ? wcount[w] : 0u;
        total += wcount[w];
    }
    if (mover) {
        const unsigned long long below = mask & ((1ull << lane) - 1ull);
        P.movers[static_cast<uint32_t>(gb) * kBlock + base + static_cast<uint32_t>(__popcll(below))] = pid;
    }
    if (threadIdx.x == 0) P.n_movers[gb] = total;
}

// --------------------------------------------------------------------------
// drift kernel
// --------------------------------------------------------------------------

// Storage index of a slot's li-th particle in the canonical order, so that
// per-workgroup reductions do not depend on the storage order: its
// cell-ordered slot's particle when the slot has a cell list (the cell-ordered
// particles of slot r sit at [off_r, off_r + n_r)), else the storage order
// (then the load order, storage being the identity until a cell list exists).
__device__ inline uint32_t canon_index(const BatchPtrs &P, const SysParams &sp, const DevScalars &sc, int li) {
    const uint32_t k = static_cast<uint32_t>(sp.off + li);
    return sc.list_valid ? P.sorted[k] : k;
}

// Kick / drift / wrap of one particle by its slot's ensemble, then its cell
// key.  Restates src/md_cuda_common.cu:381-496 (VV, NH, NPH scale+drift,
// piston) and src/md_energy_minimizer.cu:32-71 (FIRE zero/mix/drift).  Idle
// slots that keep their cell list skip the kernel; other slots that do not
// move this step (idle without a cell list, converged FIRE) still recompute
// keys so the cell list is rebuilt unchanged.
// ahead: a window's head step, whose lists must also hold through `ahead`
// lean steps: the slot rebuilds now when some particle, moving 1.5 x its
// current speed, would pass skin / 2 within them (a particle that passes it
// anyway makes its lean step fall back).
template <typename AT>
__global__ void __launch_bounds__(kBlock) k_drift(BatchPtrs P, int32_t ahead) {
    const int gb = act_block(P, xcd_block(blockIdx.x, gridDim.x));
    const int r = P.blk2sys[gb];
    const SysParams &sp = P.sp[r];
    DevScalars &sc = P.sc[r];
    const int kind = sp.kind;
    if (kind == SK_NONE && !sc.rebuild) return;   // idle slot keeping its cell list (whole workgroup)
    const int li = (gb - sp.boff) * kBlock + threadIdx.x;
    const bool reduce = (kind == SK_NVT || kind == SK_NPH);
    const Grid g = sc.g;
    double red = 0.0;
    double xi1 = 0.0;
    if (kind == SK_NVT) xi1 = nh_half_update(sc.xi, sc.K, sp);
    const bool moves = kind == SK_NVE || kind == SK_NVT || kind == SK_NPH || kind == SK_PISTON ||
                       (kind == SK_FIRE && sc.f_do_drift);
    bool mover = false;
    uint32_t mover_id = 0;
    if (li < sp.n) {
        const uint32_t i = canon_index(P, sp, sc, li);
        double2 p = P.x[i];
        const double dt = sp.dt;
        double speed2 = 0.0;   // |v| after this step's kick
        if (moves) {
            double2 q = P.v[i];
            const double2 acc = load_acc<AT>(P.a, i);
            if (kind == SK_NVE || kind == SK_PISTON) {
                q.x += 0.5 * dt * acc.x;
                q.y += 0.5 * dt * acc.y;
                p.x += dt * q.x;
                p.y += dt * q.y;
                p.x = pbc_wrap(p.x, g.Lx);
                if (kind == SK_NVE) p.y = pbc_wrap(p.y, g.Ly);
            } else if (kind == SK_NVT) {
                double vx = q.x, vy = q.y;
                vx += 0.5 * dt * (acc.x - xi1 * vx);
                vy += 0.5 * dt * (acc.y - xi1 * vy);
                q.x = vx;
                q.y = vy;
                p.x += dt * vx;
                p.y += dt * vy;
                p.x = pbc_wrap(p.x, g.Lx);
                p.y = pbc_wrap(p.y, g.Ly);
                if (sp.middle_wrap) {   // src/md_cuda_common.cu:26-65
                    double ym = fmod(p.x - sp.slab_xmin, sp.slab_period);
                    if (ym < 0.0) ym += sp.slab_period;
                    if (ym > sp.slab_width) ym = sp.slab_period - ym;
                    p.x = sp.slab_xmin + ym;
                }
            } else if (kind == SK_NPH) {
                const double lin = sc.nph_lin, vsc = sc.nph_vsc;
                q.x += 0.5 * dt * acc.x;
                q.y += 0.5 * dt * acc.y;
                q.x *= vsc;
                q.y *= vsc;
                p.x = lin * p.x + dt * q.x;
                p.y = lin * p.y + dt * q.y;
                p.x = pbc_wrap(p.x, g.Lx);
                p.y = pbc_wrap(p.y, g.Ly);
            } else {   // SK_FIRE
                if (sc.f_zero_v) {
                    q.x = 0.0;
                    q.y = 0.0;
                }
                const double alpha = sc.f_alpha;
                const double fdt = sc.f_dt;
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
                if (sp.max_step > 0.0) {
                    const double d = sqrt(ddx * ddx + ddy * ddy);
                    if (d > sp.max_step) {
                        const double f = sp.max_step / d;
                        ddx *= f;
                        ddy *= f;
                    }
                }
                p.x = pbc_wrap(p.x + ddx, g.Lx);
                p.y = pbc_wrap(p.y + ddy, g.Ly);
            }
            P.v[i] = q;
            P.x[i] = p;
            speed2 = q.x * q.x + q.y * q.y;
            if (reduce) red = 0.5 * sp.pt.m[P.cls[i]] * speed2;
        } else if (sp.wrap_keys) {
            p.x = pbc_wrap(p.x, g.Lx);
            if (g.periodic_y) p.y = pbc_wrap(p.y, g.Ly);
        }
        const int cx = cell_coord(p.x, g.Lx, g.ncx);
        const int cy = cell_coord(p.y, g.Ly, g.ncy);
        const uint32_t c = static_cast<uint32_t>(sp.coff) + static_cast<uint32_t>(cy) * g.ncx + cx;
        // movers are relative to the keys the current cell list was built
        // from: key_list after a list-mode build, else the previous step's key
        mover = !sc.list_valid || (sc.list_built ? P.key_list[i] : P.key[i]) != c;
        // displacement since the last cell-list build of a list-mode slot (x0
        // is written at every such build)
        if (sc.list_on && sc.list_built) {
            const double2 p0 = P.x0[i];
            double dx = p.x - p0.x, dy = p.y - p0.y;
            double lim2 = sp.skin_half2;
            if (kind == SK_NPH) {
                // the barostat has scaled every position by s = lin_cum since
                // the build, so two particles that were r >= r_list apart
                // are at least s r_list - |d_i| - |d_j| apart, d the
                // displacement from the scaled build position: the list holds
                // every pair inside r_c while each |d| <= (s r_list - r_c) / 2
                // (lowered by 1e-9 for rounding)
                const double sc_ = sc.lin_cum;
                dx = p.x - sc_ * p0.x;
                dy = p.y - sc_ * p0.y;
                const double a = 0.5 * (sc_ * sp.r_list - sp.rc_max) - 1.0e-9;
                lim2 = a > 0.0 ? a * a : -1.0;
            }
            if (dx > 0.5 * g.Lx) dx -= g.Lx;
            else if (dx < -0.5 * g.Lx) dx += g.Lx;
            if (dy > 0.5 * g.Ly) dy -= g.Ly;
            else if (dy < -0.5 * g.Ly) dy += g.Ly;
            double r2 = dx * dx + dy * dy;
            if (ahead > 0) {
                const double rr = sqrt(r2) + 1.5 * ahead * sp.dt * sqrt(speed2);
                r2 = rr * rr;
            }
            if (r2 > lim2) sc.rebuild = 1;   // same value from every writer
        }
        P.key[i] = c;
        mover_id = i;
    }
    append_movers(P, mover, mover_id, gb);
    if (reduce) {
        double vals[1] = {red};
        block_partials<1>(vals, 0u, P.partials, P.TB, gb);
    }
}

// --------------------------------------------------------------------------
// cell list: scan, scatter, per-cell sort + image
// --------------------------------------------------------------------------

// Per-chunk sums of count + arrivals.  A chunk lies in one slot (each slot's
// cells, its padding cell included, are a whole number of chunks); the chunks
// of a slot that does not rebuild are skipped, its cell starts stay as they are.
__global__ void __launch_bounds__(kBlock) k_scan_a(BatchPtrs P, int ncell, uint32_t *__restrict__ bsum) {
    if (!P.sc[P.cblk2sys[blockIdx.x * (kScanChunk / kBlock)]].rebuild) return;
    __shared__ uint32_t sh[kWaves];
    const int base = blockIdx.x * kScanChunk + threadIdx.x * 4;
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j)
        if (base + j < ncell) s += P.count[base + j] + P.arrivals[base + j];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kWaves; ++w) t += sh[w];
        bsum[blockIdx.x] = t;
    }
}

// Cell starts of the slots that rebuild: an exclusive scan of count +
// arrivals within the slot, from the slot's first particle.  Each slot's cells
// (its padding cell last, whose count is n_cap - n) sum to n_cap, so this
// equals the scan over the whole batch.
__global__ void __launch_bounds__(kBlock) k_scan_b(BatchPtrs P, int ncell, const uint32_t *__restrict__ bsum) {
    const int r = P.cblk2sys[blockIdx.x * (kScanChunk / kBlock)];
    const SysParams &sp = P.sp[r];
    if (!P.sc[r].rebuild) return;
    __shared__ uint32_t shw[kWaves];
    __shared__ uint32_t shp[kWaves];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    // the slot's chunks before this one
    const int b0 = sp.coff / kScanChunk;
    uint32_t pre = 0;
    for (int b = b0 + static_cast<int>(threadIdx.x); b < static_cast<int>(blockIdx.x); b += kBlock) pre += bsum[b];
    for (int o = 32; o > 0; o >>= 1) pre += __shfl_xor(pre, o, 64);
    if (lane == 0) shp[wid] = pre;
    const int base = blockIdx.x * kScanChunk + threadIdx.x * 4;
    uint32_t c[4];
    uint32_t tsum = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        c[j] = (base + j < ncell) ? P.count[base + j] + P.arrivals[base + j] : 0u;
        tsum += c[j];
    }
    uint32_t inc = tsum;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(inc, o, 64);
        if (lane >= o) inc += y;
    }
    if (lane == 63) shw[wid] = inc;
    __syncthreads();
    uint32_t chunk_pre = static_cast<uint32_t>(sp.off);
    for (int w = 0; w < kWaves; ++w) chunk_pre += shp[w];
    uint32_t wave_off = 0;
    for (int w = 0; w < wid; ++w) wave_off += shw[w];
    uint32_t run = chunk_pre + wave_off + inc - tsum;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (base + j < ncell) P.cs[base + j] = run;
        run += c[j];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == kBlock - 1) P.cs[ncell] = run;
}

// Drift reduction of a slot (Nose-Hoover / NPH kinetic energy after the
// first half kick), combined from the drift workgroups' partials by one whole
// workgroup: K_half, and for Nose-Hoover the mid-step xi.
__device__ inline void finalize_drift(const BatchPtrs &P, const SysParams &sp, DevScalars &sc) {
    if (sp.n <= 0 || sp.nblk <= 0 || (sp.kind != SK_NVT && sp.kind != SK_NPH)) return;
    double out[1];
    slot_reduce<1>(0u, P.partials, P.TB, sp.boff, act_blocks(sp), out);
    if (threadIdx.x == 0) {
        sc.K_half = out[0];
        if (sp.kind == SK_NVT) {
            const double xi1 = nh_half_update(sc.xi, sc.K, sp);
            sc.xi_mid[0] = nh_half_update(xi1, out[0], sp);
            sc.eta += 0.5 * sp.dt * (xi1 + sc.xi_mid[0]);
        }
    }
}

// cell-list entries a per-cell thread loads at a time
constexpr int kCellUnroll = 8;

// Workgroups [0, ncb): one thread per cell of the new grid (and of the
// previous grid when it was larger): count[c] = particles of the previous list
// of cell c that stay in c; cells beyond the new grid get zero.  Workgroups
// [ncb, ncb + TB): drift workgroup b's movers take an arrival index in their
// new cell (atomics on distinct cells).  The two parts touch different arrays.
__global__ void __launch_bounds__(kBlock) k_stay_arrive(BatchPtrs P, int ncb) {
    if (static_cast<int>(blockIdx.x) >= ncb) {
        const uint32_t b = static_cast<uint32_t>(act_block(P, blockIdx.x - ncb));   // k_drift's workgroup
        if (!P.sc[P.blk2sys[b]].rebuild) return;
        const uint32_t nm = P.n_movers[b];
        if (threadIdx.x < nm) {
            const uint32_t m = b * kBlock + threadIdx.x;
            P.mover_off[m] = atomicAdd(&P.arrivals[P.key[P.movers[m]]], 1u);
        }
        return;
    }
    const int gb = act_cblock(P, xcd_block(blockIdx.x, ncb));
    const int r = P.cblk2sys[gb];
    const SysParams &sp = P.sp[r];
    DevScalars &sc = P.sc[r];
    const int lc = (gb - sp.cboff) * kBlock + threadIdx.x;
    // a slot that keeps its lists keeps count = cell occupancy (set by
    // k_cell_sort), so the scan reproduces its cell starts
    if (!sc.rebuild) return;
    const int n_new = sc.g.ncx * sc.g.ncy;
    const int n_old = sc.list_valid ? sc.g_old.ncx * sc.g_old.ncy : 0;
    // without a previous list (new or reused slot) every cell of the slot is
    // cleared, whatever an earlier occupant left in it
    const int limit = sc.list_valid ? (n_new > n_old ? n_new : n_old) : sp.ccap;
    if (lc >= limit || lc >= sp.ccap) return;
    const uint32_t c = static_cast<uint32_t>(sp.coff + lc);
    uint32_t stay = 0;
    if (sp.n > 0 && lc < n_old && lc < n_new) {
        const uint32_t s0 = P.cs_old[c], s1 = P.cs_old[c + 1];
        // eight entries at a time, their loads in flight together
        for (uint32_t k = s0; k < s1; k += kCellUnroll) {
            uint32_t id[kCellUnroll];
#pragma unroll
            for (int u = 0; u < kCellUnroll; ++u) id[u] = k + u < s1 ? P.sorted_old[k + u] : 0u;
#pragma unroll
            for (int u = 0; u < kCellUnroll; ++u) stay += (k + u < s1 && P.key[id[u]] == c) ? 1u : 0u;
        }
    }
    P.count[c] = stay;
}

// Workgroups [0, ncb): one thread per cell of the new grid copies the cell's
// stayers in their previous order (ascending id) to the start of the cell.
// Workgroups [ncb, ncb + TB): drift workgroup b's movers go after the stayers
// of their new cell, at their arrival index.
__global__ void __launch_bounds__(kBlock) k_fill(BatchPtrs P, int ncb) {
    if (static_cast<int>(blockIdx.x) >= ncb) {
        const uint32_t b = static_cast<uint32_t>(act_block(P, blockIdx.x - ncb));   // k_drift's workgroup
        if (!P.sc[P.blk2sys[b]].rebuild) return;
        const uint32_t nm = P.n_movers[b];
        if (threadIdx.x < nm) {
            const uint32_t m = b * kBlock + threadIdx.x;
            const uint32_t id = P.movers[m];
            const uint32_t c = P.key[id];
            P.sorted[P.cs[c] + P.count[c] + P.mover_off[m]] = id;
        }
        return;
    }
    const int gb = act_cblock(P, xcd_block(blockIdx.x, ncb));
    const int r = P.cblk2sys[gb];
    const SysParams &sp = P.sp[r];
    const DevScalars &sc = P.sc[r];
    if (!sc.rebuild) return;
    const int lc = (gb - sp.cboff) * kBlock + threadIdx.x;
    const int n_new = sc.g.ncx * sc.g.ncy;
    const int n_old = sc.list_valid ? sc.g_old.ncx * sc.g_old.ncy : 0;
    if (sp.n > 0 && lc < n_new && lc < n_old) {
        const uint32_t c = static_cast<uint32_t>(sp.coff + lc);
        uint32_t dst = P.cs[c];
        const uint32_t s0 = P.cs_old[c], s1 = P.cs_old[c + 1];
        for (uint32_t k = s0; k < s1; k += kCellUnroll) {
            uint32_t id[kCellUnroll];
            bool st[kCellUnroll];
#pragma unroll
            for (int u = 0; u < kCellUnroll; ++u) id[u] = k + u < s1 ? P.sorted_old[k + u] : 0u;
#pragma unroll
            for (int u = 0; u < kCellUnroll; ++u) st[u] = k + u < s1 && P.key[id[u]] == c;
#pragma unroll
            for (int u = 0; u < kCellUnroll; ++u)
                if (st[u]) P.sorted[dst++] = id[u];
        }
    }
}

// Sort of the n <= N storage indices at a by their original index (one
// thread): the original indices loaded together into registers (padded with
// the largest key), a bitonic network with compile-time indices, stored back
// as storage indices.
template <int N>
__device__ inline void sort_cell(uint32_t *a, uint32_t n, const uint32_t *pid, const uint32_t *sid) {
    uint32_t v[N];
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = static_cast<uint32_t>(i) < n ? a[i] : 0u;
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = static_cast<uint32_t>(i) < n ? pid[v[i]] : 0xFFFFFFFFu;
#pragma unroll
    for (int k = 2; k <= N; k <<= 1)
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
            for (int i = 0; i < N; ++i) {
                const int l = i ^ j;
                if (l > i) {
                    const uint32_t x = v[i], y = v[l];
                    const uint32_t lo = x < y ? x : y, hi = x < y ? y : x;
                    const bool up = (i & k) == 0;
                    v[i] = up ? lo : hi;
                    v[l] = up ? hi : lo;
                }
            }
#pragma unroll
    for (int i = 0; i < N; ++i)
        if (static_cast<uint32_t>(i) < n) a[i] = sid[v[i]];
}

// One thread per cell: cells that received movers restore the order of
// ascending original index (stayers are in order, the scatter's atomics leave
// the arrivals in any order); every cell records its range for the next
// step's update.
__global__ void __launch_bounds__(kBlock) k_cell_sort(BatchPtrs P) {
    const int gb = act_cblock(P, xcd_block(blockIdx.x, gridDim.x));
    const int r = P.cblk2sys[gb];
    const SysParams &sp = P.sp[r];
    const Grid g = P.sc[r].g;
    const int lc = (gb - sp.cboff) * kBlock + threadIdx.x;
    const int ncell = g.ncx * g.ncy;
    if (sp.n <= 0 || lc >= ncell || !P.sc[r].rebuild) return;
    if (lc == 0) {
        P.sc[r].list_valid = 1;
        P.sc[r].list_built = 0;   // built from this step's keys; k_list_build marks a list-mode build
    }
    const int c = sp.coff + lc;
    const uint32_t s0 = P.cs[c], s1 = P.cs[c + 1];
    P.cs_old[c] = s0;
    if (lc == ncell - 1) P.cs_old[c + 1] = s1;
    P.count[c] = s1 - s0;   // occupancy, for steps that keep this list
    if (P.arrivals[c] == 0) return;
    P.arrivals[c] = 0;
    uint32_t *sorted = P.sorted;
    const uint32_t n = s1 - s0;
    if (n <= 16u) {
        sort_cell<16>(sorted + s0, n, P.pid, P.sid);
        return;
    }
    if (n <= 32u) {
        sort_cell<32>(sorted + s0, n, P.pid, P.sid);
        return;
    }
    for (uint32_t k = s0 + 1; k < s1; ++k) {   // larger cells: insertion sort in place
        const uint32_t id = sorted[k];
        const uint32_t key = P.pid[id];
        uint32_t j = k;
        while (j > s0) {
            const uint32_t prev = sorted[j - 1];
            if (P.pid[prev] <= key) break;
            sorted[j] = prev;
            --j;
        }
        sorted[j] = id;
    }
}

// One thread per cell-ordered slot: the particle's image (FP32 position
// relative to its cell centre, or the wrapped FP64 position), class, cell,
// and its id for the next step's update.  Writes are coalesced.
template <typename IT>
__global__ void __launch_bounds__(kBlock) k_slot_image(BatchPtrs P) {
    const int gb = act_block(P, xcd_block(blockIdx.x, gridDim.x));
    const int r = P.blk2sys[gb];
    const SysParams &sp = P.sp[r];
    DevScalars &sc = P.sc[r];
    if (gb == sp.boff) finalize_drift(P, sp, sc);   // the slot's first workgroup: K_half, xi_mid
    const int li = (gb - sp.boff) * kBlock + threadIdx.x;
    if (li >= sp.n) return;
    const Grid g = sc.g;
    const bool reb = sc.rebuild;
    const uint32_t k = P.cs[sp.coff] + static_cast<uint32_t>(li);
    const uint32_t id = P.sorted[k];
    // the particle's cell in the list: the current one on a rebuild, else the
    // one of the last build (the particle may have left it since)
    const uint32_t lc = reb ? P.key[id] - static_cast<uint32_t>(sp.coff) : P.scell[k];
    double2 p = P.x[id];
    if (sp.wrap_keys) {
        p.x = pbc_wrap(p.x, g.Lx);
        if (g.periodic_y) p.y = pbc_wrap(p.y, g.Ly);
    }
    IT *img = static_cast<IT *>(P.img);
    if constexpr (std::is_same_v<IT, int2>) {
        const uint32_t cy = lc / static_cast<uint32_t>(g.ncx);
        const uint32_t cx = lc - cy * static_cast<uint32_t>(g.ncx);
        double dx = p.x - (static_cast<double>(cx) + 0.5) * g.w;
        double dy = p.y - (static_cast<double>(cy) + 0.5) * g.h;
        if (!reb) {   // minimum image: the particle may have crossed the box edge
            if (dx > 0.5 * g.Lx) dx -= g.Lx;
            else if (dx < -0.5 * g.Lx) dx += g.Lx;
            if (dy > 0.5 * g.Ly) dy -= g.Ly;
            else if (dy < -0.5 * g.Ly) dy += g.Ly;
        }
        const int2 im = make_int2(to_fixed(dx, g.qs), to_fixed(dy, g.qs));
        img[k] = im;
        if (reb && sp.list_mode) P.img0[k] = im;   // build-time image, for the lean steps' check
    } else {
        img[k] = p;
    }
    if (reb) {
        P.icls[k] = P.cls[id];
        P.scell[k] = lc;
        P.sorted_old[k] = id;
        if (sp.list_mode) {
            P.x0[id] = p;
            P.key_list[id] = P.key[id];
            if (li == 0) sc.lin_cum = 1.0;   // no barostat scaling since this list build
        }
    }
}

// Lean step of a rebuild window (list mode: NVE / NVT slots, idle slots),
// with step parity `parity`: the first kick, drift and wrap of k_drift and
// the image of k_slot_image in one pass over the cell-ordered slots of the
// last cell list, and the scalar work a step starts with, so that a lean step
// is this kernel and the pair pass.  No keys or movers (no cell-list kernels
// run this step).
//
// Each workgroup (kLeanThreads threads, kLeanSpan slots) of a Nose-Hoover
// slot first combines the previous pair pass's partials itself, in
// slot_reduce's fixed order, so every workgroup of the slot gets the same K,
// sum m v.a and sum m |a|^2 / 2 with no synchronisation between workgroups.
// From them: xi1 = the first half update of xi (the kick's friction), and,
// since the kick is v' = c v + (dt/2) a with c = 1 - dt xi1 / 2, K(v') = c^2 K
// + (c dt / 2) sum m v.a + (dt^2 / 4) sum m |a|^2 / 2 and the mid-step xi for
// the pair pass's second kick.  The slot's first workgroup stores them:
// xi_mid[parity] (the other entry, the previous step's, is what every
// workgroup reads), the pass's sums as finalize_pair would, and it clears
// fallback[parity ^ 1] for the next lean step.
//
// A particle beyond skin / 2 of its build position sets fallback[parity]: the
// step's pair pass takes the 5x5 cells of the last cell list; beyond (2 min(w,
// h) - rc_max) / 2 the step is flagged as not exact.
template <typename AT>
__global__ void __launch_bounds__(kLeanThreads) k_lean(BatchPtrs P, int32_t parity) {
    const int gb = xcd_block(blockIdx.x, gridDim.x);
    const int r = P.lb2sys[gb];
    const SysParams &sp = P.sp[r];
    DevScalars &sc = P.sc[r];
    const int kind = sp.kind;
    const bool moves = kind == SK_NVE || kind == SK_NVT;
    const bool first = gb == sp.lboff && threadIdx.x == 0;
    const uint32_t k0 = P.cs[sp.coff];
    const int base = (gb - sp.lboff) * kLeanSpan + static_cast<int>(threadIdx.x);
    // the particles' data first: their loads are in flight during the reduction
    bool act[kLeanPer];
    uint32_t ids[kLeanPer], lcs[kLeanPer];
    double2 xs[kLeanPer], vs[kLeanPer], as[kLeanPer];
#pragma unroll
    for (int j = 0; j < kLeanPer; ++j) {
        const int li = base + j * kLeanThreads;
        act[j] = li < sp.n;
        if (act[j]) {
            const uint32_t k = k0 + static_cast<uint32_t>(li);
            ids[j] = P.sorted[k];
            lcs[j] = P.scell[k];
        }
    }
#pragma unroll
    for (int j = 0; j < kLeanPer; ++j) {
        if (act[j]) {
            xs[j] = P.x[ids[j]];
            if (moves) {
                vs[j] = P.v[ids[j]];
                as[j] = load_acc<AT>(P.a, ids[j]);
            }
        }
    }
    // the grid and the build images too (after the stores below, the
    // compiler could not move these loads ahead of them)
    const Grid g = sc.g;
    const bool check = moves && sc.list_valid;
    int2 i0s[kLeanPer];
#pragma unroll
    for (int j = 0; j < kLeanPer; ++j)
        if (act[j] && check) i0s[j] = P.img0[k0 + static_cast<uint32_t>(base + j * kLeanThreads)];
    double xi1 = 0.0;
    if (kind == SK_NVT && sp.nblk > 0) {   // uniform per workgroup
        double out[3];
        slot_reduce<3>(0u, P.partials, P.TB, sp.boff, act_blocks(sp), out);
        const double xi = sc.xi_mid[parity ^ 1];
        xi1 = nh_half_update(xi, out[0], sp);
        if (first) {
            const double c = 1.0 - 0.5 * sp.dt * xi1;
            const double Kh = c * c * out[0] + 0.5 * c * sp.dt * out[1] + 0.25 * sp.dt * sp.dt * out[2];
            sc.K = out[0];
            sc.S_va = out[1];
            sc.S_aa = out[2];
            sc.xi = xi;
            sc.K_half = Kh;
            sc.xi_mid[parity] = nh_half_update(xi1, Kh, sp);
            sc.eta += 0.5 * sp.dt * (xi1 + sc.xi_mid[parity]);
            sc.pair_pending = 0;
        }
    }
    if (first) sc.fallback[parity ^ 1] = 0;
    // thresholds in fixed-point units squared (q2f is a power of two)
    const float thr = static_cast<float>(sp.skin_half2) * (1.0f - 1.0e-5f) * g.q2f;
    const float reach = 0.5f * static_cast<float>(2.0 * fmin(g.w, g.h) - sp.rc_max);
    const float thr_v = reach * reach * (1.0f - 1.0e-5f) * g.q2f;
    bool fb = false, bad = false;
#pragma unroll
    for (int j = 0; j < kLeanPer; ++j) {
        if (!act[j]) continue;
        const uint32_t k = k0 + static_cast<uint32_t>(base + j * kLeanThreads);
        const uint32_t id = ids[j];
        double2 p = xs[j];
        if (moves) {
            const double dt = sp.dt;
            double2 q = vs[j];
            const double2 acc = as[j];
            if (kind == SK_NVE) {   // src/md_cuda_common.cu:381-394
                q.x += 0.5 * dt * acc.x;
                q.y += 0.5 * dt * acc.y;
                p.x += dt * q.x;
                p.y += dt * q.y;
            } else {                // src/md_cuda_common.cu:454-478
                double vx = q.x, vy = q.y;
                vx += 0.5 * dt * (acc.x - xi1 * vx);
                vy += 0.5 * dt * (acc.y - xi1 * vy);
                q.x = vx;
                q.y = vy;
                p.x += dt * vx;
                p.y += dt * vy;
            }
            p.x = pbc_wrap(p.x, g.Lx);
            p.y = pbc_wrap(p.y, g.Ly);
            if (sp.middle_wrap) {   // do_middle_wrap: as in k_drift (src/md_cuda_common.cu:26-65)
                double ym = fmod(p.x - sp.slab_xmin, sp.slab_period);
                if (ym < 0.0) ym += sp.slab_period;
                if (ym > sp.slab_width) ym = sp.slab_period - ym;
                p.x = sp.slab_xmin + ym;
            }
            P.v[id] = q;
            P.x[id] = p;
        }
        // image relative to the centre of the particle's cell in the last cell
        // list, minimum image (the particle may have crossed the box edge since)
        const uint32_t lc = lcs[j];
        const uint32_t cy = lc / static_cast<uint32_t>(g.ncx);
        const uint32_t cx = lc - cy * static_cast<uint32_t>(g.ncx);
        double dx = p.x - (static_cast<double>(cx) + 0.5) * g.w;
        double dy = p.y - (static_cast<double>(cy) + 0.5) * g.h;
        if (dx > 0.5 * g.Lx) dx -= g.Lx;
        else if (dx < -0.5 * g.Lx) dx += g.Lx;
        if (dy > 0.5 * g.Ly) dy -= g.Ly;
        else if (dy < -0.5 * g.Ly) dy += g.Ly;
        const int2 im = make_int2(to_fixed(dx, g.qs), to_fixed(dy, g.qs));
        static_cast<int2 *>(P.img)[k] = im;
        if (check) {
            // displacement since the build: the two images share the build
            // cell's centre (exact difference, squared in FP32; the
            // thresholds are lowered by 1e-5 relative so that a rounding near
            // them errs towards the fallback)
            const int2 i0 = i0s[j];
            const float ex = static_cast<float>(im.x - i0.x), ey = static_cast<float>(im.y - i0.y);
            const float d2 = __builtin_fmaf(ex, ex, ey * ey);
            fb = fb || d2 > thr;
            bad = bad || d2 > thr_v;
        }
    }
    if (fb) sc.fallback[parity] = 1;   // same value from every writer
    if (bad) sc.violated = 1;
}

// Neighbour-list entry (list mode): (partner slot << 3), so that entry &
// kListOff is the byte offset of the partner's fixed-point image; bit 2 set when the
// partner has the particle's own type class; the neighbour cell's offset
// (ox, oy), -1..1 each, as the code c = (ox + 1) + 3 (oy + 1) in 0..8, whose
// bits 0-1 sit in entry bits 30-31 and bits 2-3 in entry bits 0-1: rotated
// left by 13 the entry holds c in bits 11-14, the byte offset of row c of a
// [9][kBlock] table of int2 (list_shift_off).  Partner slots are below 2^27.
constexpr uint32_t kListOff = 0x3FFFFFF8u;
constexpr uint32_t kListSame = 4u;
// list rows a lane stages in LDS per pass of the list build
constexpr int kListStage = 32;

__device__ inline uint32_t list_entry(uint32_t k, int ox, int oy, bool same) {
    const uint32_t c = static_cast<uint32_t>((ox + 1) + 3 * (oy + 1));
    return (k << 3) | (same ? kListSame : 0u) | (c >> 2) | ((c & 3u) << 30);
}
__device__ inline uint32_t list_code(uint32_t v) { return (v >> 30) | ((v & 3u) << 2); }
// byte offset of the entry's row in a [9][kBlock] int2 table
__device__ inline uint32_t list_shift_off(uint32_t v) { return __builtin_amdgcn_alignbit(v, v, 19) & 0x7800u; }
static_assert(kBlock * sizeof(int2) == 0x800u, "list_shift_off assumes 2048-byte table rows");
// the offset's components from the code: oy + 1 = c / 3 (as (11 c) >> 5 for
// c <= 8), ox + 1 = c - 3 (oy + 1)
__device__ inline int list_oyi(uint32_t v) { return static_cast<int>((list_code(v) * 11u) >> 5) - 1; }
__device__ inline int list_oxi(uint32_t v) {
    const uint32_t c = list_code(v);
    return static_cast<int>(c - 3u * ((c * 11u) >> 5)) - 1;
}
// Padding entry of slot s: the particle itself one cell to the side, at
// distance w >= r_list > r_c, so it is never inside the cutoff.
__device__ inline uint32_t list_pad(uint32_t s) { return list_entry(s, 1, 0, true); }
// ncount[s] = rows | kListGuard when some pair of the wave's lists may come
// closer than the FP64 guard distance before the next rebuild, | kListMixed
// when some entry of the wave's lists pairs two type classes (else every
// lane's partners have the lane's own class, and the pair pass takes the
// lane's own-class constants without a per-entry select)
constexpr uint32_t kListGuard = 0x80000000u;
constexpr uint32_t kListMixed = 0x40000000u;
constexpr uint32_t kListRows = ~(kListGuard | kListMixed);

// Neighbour list of every cell-ordered slot of the slots that rebuild in list
// mode: the partners within r_list in the 3x3 neighbour cells, in the pair
// kernel's order (rows -1, 0, +1; cells -1, 0, +1; ascending id).  Entry e of
// slot s is nlist[e * N + s].  Every lane of a wave (64 consecutive slots,
// the pair kernel's wave) gets the same length: the wave's longest list
// rounded up to a multiple of 4, the rows past a lane's own list holding
// list_pad.  A lane stages its entries in its own column of LDS and the wave
// writes whole rows, so the stores are coalesced; the class bit is added as
// the rows are written (one class load per entry kept rather than per
// candidate); lists longer than kListStage take one more scan per kListStage
// rows.  Between rebuilds no
// particle moves more than skin / 2 from its build position, so a pair comes
// closer than the FP64 guard distance only if it is closer than
// guard + skin now; the wave's length carries kListGuard when one is.
__global__ void __launch_bounds__(kBlock) k_list_build(BatchPtrs P) {
    const int gb = act_block(P, xcd_block(blockIdx.x, gridDim.x));
    const int r = P.blk2sys[gb];
    const SysParams &sp = P.sp[r];
    DevScalars &sc = P.sc[r];
    if (!sc.list_on || !sc.rebuild || sp.n <= 0) return;
    __shared__ uint32_t stage[kListStage + 1][kBlock];   // + a spare row for rejected candidates
    const int tid = static_cast<int>(threadIdx.x);
    const int li = (gb - sp.boff) * kBlock + tid;
    const bool active = li < sp.n;
    const Grid g = sc.g;
    const int2 *img = static_cast<const int2 *>(P.img);
    const char *imgb = static_cast<const char *>(P.img);
    // r^2 thresholds in fixed-point units squared
    const float rl2 = static_cast<float>(sp.r_list * sp.r_list) * g.q2f;
    const uint32_t cap = static_cast<uint32_t>(sp.list_cap);   // a multiple of 4
    const size_t stride = static_cast<size_t>(P.N);
    uint32_t s = 0;
    int cx = 0, cy = 0, ci = 0;
    int2 pi = make_int2(0, 0);
    if (active) {
        s = P.cs[sp.coff] + static_cast<uint32_t>(li);
        const uint32_t lc = P.scell[s];
        cx = static_cast<int>(lc % static_cast<uint32_t>(g.ncx));
        cy = static_cast<int>(lc / static_cast<uint32_t>(g.ncx));
        pi = img[s];
        ci = P.icls[s];
    }
    // pairs closer than this may reach the guard distance before a rebuild
    // (with a margin for the FP32 difference between this and the pair form)
    const float rgd = sqrtf(sp.pt.r2_guard) + 2.0f * sqrtf(static_cast<float>(sp.skin_half2)) + 1.0e-3f;
    const float rg2 = rgd * rgd * g.q2f;
    // the smallest r^2 of any candidate (as its IEEE bit pattern less one)
    uint32_t near_rb = 0xFFFFFFFFu;
    bool mixed = false;
    uint32_t rows = 0, cnt = 0;
    for (uint32_t base = 0;; base += kListStage) {
        cnt = 0;
        if (active) {
            for (int oy = -1; oy <= 1; ++oy) {
                int yy = cy + oy;
                if (yy < 0) yy += g.ncy;
                else if (yy >= g.ncy) yy -= g.ncy;
                const int rowbase = sp.coff + yy * g.ncx;
                const int piy = pi.y - oy * g.hq;
                // the row's three cells (periodic in x) as one run of candidates:
                // cell -1 [a0, a0 + na), cell 0 [b0, ...), cell +1 [c0, ...)
                const int xl = cx == 0 ? g.ncx - 1 : cx - 1;
                const int xr = cx == g.ncx - 1 ? 0 : cx + 1;
                const uint32_t a0 = P.cs[rowbase + xl], a1 = P.cs[rowbase + xl + 1];
                const uint32_t b0 = P.cs[rowbase + cx], b1 = P.cs[rowbase + cx + 1];
                const uint32_t c0 = P.cs[rowbase + xr], c1 = P.cs[rowbase + xr + 1];
                const uint32_t na = a1 - a0, nab = na + (b1 - b0), nt = nab + (c1 - c0);
                // candidate t of the run is slot t + off(t), off = a0, b0 - na
                // or c0 - nab by the cell it falls in; its entry without the
                // class bit is (t << 3) + ((off << 3) + code bits) (mod 2^32:
                // slots are below 2^27, so the slot field neither wraps nor
                // reaches the code bits), its image offset the entry's slot
                // field.  Per cell also the particle's image moved by minus the
                // cell's offset (selects, no branches).  Past the run: the
                // particle itself one cell to the side, at distance w >= r_list,
                // never accepted.
                const uint32_t e1 = (a0 << 3) + list_entry(0u, -1, oy, false);
                const uint32_t e2 = ((b0 - na) << 3) + list_entry(0u, 0, oy, false);
                const uint32_t e3 = ((c0 - nab) << 3) + list_entry(0u, 1, oy, false);
                const uint32_t epast = list_entry(s, 1, oy, false);
                const int pxa = pi.x + g.wq, pxc = pi.x - g.wq;
                // four candidates per pass, their loads in flight together
                for (uint32_t t = 0; t < nt; t += 4) {
                    uint32_t en[4];
                    int px[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const uint32_t tt = t + static_cast<uint32_t>(u);
                        const bool g1 = tt >= na, g2 = tt >= nab;
                        en[u] = tt < nt ? (tt << 3) + (g2 ? e3 : (g1 ? e2 : e1)) : epast;
                        px[u] = g2 ? pxc : (g1 ? pi.x : pxa);
                    }
                    int2 pj[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) pj[u] = *reinterpret_cast<const int2 *>(imgb + (en[u] & kListOff));
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const float dx = static_cast<float>(px[u] - pj[u].x);
                        const float dy = static_cast<float>(piy - pj[u].y);
                        const float rr = __builtin_fmaf(dx, dx, dy * dy);
                        // 0 < r^2 (itself, and coincident particles, as in the
                        // cell pass, are not partners) and r^2 < r_list^2, as
                        // IEEE bit patterns
                        const uint32_t rb = __float_as_uint(rr) - 1u;
                        const bool acc = rb < __float_as_uint(rl2) - 1u;
                        near_rb = rb < near_rb ? rb : near_rb;
                        // unconditional store: a rejected candidate, or an entry
                        // outside this pass's rows, goes to the spare row
                        const uint32_t rel = cnt - base;
                        const uint32_t row = acc ? (rel < static_cast<uint32_t>(kListStage) ? rel
                                                                                           : static_cast<uint32_t>(kListStage))
                                                 : static_cast<uint32_t>(kListStage);
                        stage[row][tid] = en[u];
                        cnt += acc ? 1u : 0u;
                    }
                }
            }
        }
        if (base == 0) {   // the wave's row count
            uint32_t m = active ? cnt : 0u;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                const uint32_t t = static_cast<uint32_t>(__shfl_xor(static_cast<int>(m), o, 64));
                m = t > m ? t : m;
            }
            m = (m + 3u) & ~3u;
            rows = m < cap ? m : cap;
        }
        const uint32_t end = rows < base + kListStage ? rows : base + kListStage;
        if (active) {
            // the rows of this pass (four at a time: rows, base and end are
            // multiples of 4), each entry with its class bit from the
            // partner's class; rows past the list hold the padding entry
            const uint32_t pad = list_pad(s);
            for (uint32_t e = base; e < end; e += 4) {
                uint32_t v[4];
                int cj[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const uint32_t ee = e + static_cast<uint32_t>(u);
                    v[u] = ee < cnt ? stage[ee - base][tid] : pad;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    cj[u] = e + static_cast<uint32_t>(u) < cnt ? P.icls[(v[u] & kListOff) >> 3] : ci;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const uint32_t ee = e + static_cast<uint32_t>(u);
                    if (ee < cnt) {
                        const bool same = cj[u] == ci;
                        v[u] |= same ? kListSame : 0u;
                        mixed = mixed || !same;
                    }
                    P.nlist[static_cast<size_t>(ee) * stride + s] = v[u];
                }
            }
        }
        if (base + kListStage >= rows) break;
    }
    const bool guard = __any(near_rb < __float_as_uint(rg2) - 1u), mix = __any(mixed);
    if (active) {
        P.ncount[s] = rows | (guard ? kListGuard : 0u) | (mix ? kListMixed : 0u);
        if (cnt > cap) sc.list_overflow = 1;   // this step's pair pass uses the cells instead
    }
    if (li == 0) sc.list_built = 1;
}

// Particle counts in n_bins slabs along x of slot r, type class 0 (A) in
// counts[0, nb), every other type (B) in counts[nb, 2 nb): the reference's
// local_density_profile_kernel (src/md_cuda_common.cu:827-877) on one domain
// [xmin, xmax] = [0, Lx] with the current box, bin = int((x - xmin) / (xmax -
// xmin) * nb) clamped to [0, nb), particles outside [xmin, xmax] not
// counted.  Integer counts: the result does not depend on the order of the
// atomics.  Dynamic LDS: 2 nb counters.
__global__ void __launch_bounds__(kBlock) k_n_profile(BatchPtrs P, int32_t r, int32_t nb, int32_t *counts) {
    extern __shared__ int32_t hist[];
    const SysParams &sp = P.sp[r];
    for (int j = threadIdx.x; j < 2 * nb; j += kBlock) hist[j] = 0;
    __syncthreads();
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i < sp.n) {
        const int64_t k = sp.off + i;
        const double x = P.x[k].x;
        const double xmin = 0.0, xmax = P.sc[r].g.Lx;
        if (x >= xmin && x <= xmax) {
            int bin = static_cast<int>((x - xmin) / (xmax - xmin) * nb);
            bin = bin < 0 ? 0 : (bin >= nb ? nb - 1 : bin);
            atomicAdd(&hist[(P.cls[k] == 0 ? 0 : nb) + bin], 1);
        }
    }
    __syncthreads();
    for (int j = threadIdx.x; j < 2 * nb; j += kBlock)
        if (hist[j] != 0) atomicAdd(&counts[j], hist[j]);
}

// --------------------------------------------------------------------------
// pair kernel
// --------------------------------------------------------------------------

enum Acc { ACC_COMP32 = 0, ACC_FP64 = 1 };

// The sums a pair pass forms besides the forces (its OBS template argument),
// for the slots that take them (energy passes, NPH steps):
//   SUMS_ENERGY  U, the virial W and dU/dlambda, each pair added in FP64 (energy
//                passes, and the FP64 form's NPH steps);
//   SUMS_VIRIAL  W alone, for the NPH steps of the FP32 forms: the pair's r.F
//                is its force factor times r^2, summed like its force (FP32
//                runs entering an FP64 sum in the compensated form, FP64 per
//                pair in the FP64-accumulating form); FP64-evaluated pairs add
//                theirs in FP64.
enum PairSums { SUMS_NONE = 0, SUMS_ENERGY = 1, SUMS_VIRIAL = 2 };

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
template <int OBS>
__device__ inline bool pair_fp64(double dx, double dy, int pidx, const PairTable &pt, double &fx, double &fy,
                                 ObsSum &ob, bool want_obs) {
    const double r2 = dx * dx + dy * dy;
    if (!(r2 > 0.0 && r2 < pt.rc2[pidx])) return false;
    const double r2i = 1.0 / r2;
    const double sr2 = pt.sig2[pidx] * r2i;
    const double sr6 = sr2 * sr2 * sr2;
    const double sr12 = sr6 * sr6;
    const double tmp = pt.eps24[pidx] * (2.0 * sr12 - sr6) * r2i;
    fx = tmp * dx;
    fy = tmp * dy;
    if constexpr (OBS == SUMS_ENERGY) {
        if (want_obs) {
            ob.u += 0.5 * (pt.eps4[pidx] * (sr12 - sr6) - pt.vc[pidx]);
            ob.w += 0.5 * (dx * fx + dy * fy);
            ob.lam += 0.5 * (fx * dx - fy * dy);
        }
    } else if constexpr (OBS == SUMS_VIRIAL) {
        if (want_obs) ob.w += 0.5 * (dx * fx + dy * fy);
    }
    return true;
}

// Last-workgroup finalisation of a slot's pair-kernel reduction out[0..2].
__device__ inline void finalize_pair(int pmode, const double (&out)[3], const SysParams &sp, const Grid &g,
                                     DevScalars &sc) {
    switch (pmode) {
    case PM_OBS:
        sc.U = out[0];
        sc.W = out[1];
        sc.lam = out[2];
        break;
    case PM_WALLS:   // cal_forces_piston, src/md_env.cu:1570-1573
        sc.F_p = out[1];
        sc.wallE = out[2];
        sc.last_P = (g.Lx > 0.0) ? sc.F_p / g.Lx : 0.0;
        break;
    case PM_NVT:
        sc.K = out[0];
        sc.S_va = out[1];
        sc.S_aa = out[2];
        sc.xi = sc.xi_mid[sc.pair_parity];
        break;
    case PM_NPH: {   // src/md_env.cu:1896-1917
        const double W = out[1];
        const double area = g.Lx * g.Ly;
        const double P_new = (area > 0.0) ? (2.0 * sc.K_half + W) / (2.0 * area) : 0.0;
        double pA = sc.p_A + 0.5 * sp.dt * (P_new - sp.p_ext);
        if (sc.clamped) {
            if (g.Ly <= sc.Ly_floor && pA < 0.0) pA = 0.0;
            if (g.Ly >= sc.Ly_ceil && pA > 0.0) pA = 0.0;
        }
        sc.p_A = pA;
        sc.last_P = P_new;
        sc.K = out[0];
        sc.W = W;
        break;
    }
    case PM_PISTON: {   // src/md_env.cu:2004-2022
        sc.F_p = out[1];
        sc.wallE = out[2];
        const double acc_new = (sc.F_p - sp.p_ext * g.Lx) / sp.W_baro;
        double vp = sc.v_p + 0.5 * sp.dt * acc_new;
        if (sc.clamped) {
            if (g.Ly <= sc.Ly_floor && vp < 0.0) vp = 0.0;
            if (g.Ly >= sc.Ly_ceil && vp > 0.0) vp = 0.0;
        }
        sc.v_p = vp;
        sc.last_P = sc.F_p / g.Lx;
        break;
    }
    case PM_FIRE: {   // src/md_energy_minimizer.cu:170-203
        const double maxF = out[0], power = out[1];
        sc.f_maxF = maxF;
        sc.f_power = power;
        const int step = sc.f_step;
        if (maxF < sp.ftol) {
            sc.f_done = 1;
            sc.f_conv_step = step;
        } else if (power > 0.0) {
            sc.f_streak += 1;
            if (sc.f_streak > sp.N_min) {
                const double inc = sc.f_dt * sp.f_inc;
                sc.f_dt = (sp.dt_max < inc) ? sp.dt_max : inc;
                const double al = sc.f_alpha * 0.99;
                sc.f_alpha = (0.0 < al) ? al : 0.0;
            }
            sc.f_zero_v = 0;
        } else {
            sc.f_dt *= sp.f_dec;
            sc.f_streak = 0;
            sc.f_alpha = sp.alpha0;
            sc.f_zero_v = 1;
        }
        sc.f_step = step + 1;
        break;
    }
    default:
        break;
    }
}

// Per-particle constants of the FP32 pair path: "s" entries for partners of
// the particle's own class, "o" for the AB combination.  A displacement
// arrives in fixed-point units (times 2^qk, Grid), so squared distances
// (rc2, r2_guard, rcore2) are compared in those units squared, 1/r^2 is
// brought back to units of sigma^-2 by q2 = 2^(2 qk), and the force factors
// e24, fa, fb carry 2^-qk, so that their product with the fixed-point
// displacement is the force.  All scalings are powers of two (exact).
struct PairF32 {
    float rc2s, rc2o, sig2s, sig2o, e24s, e24o, e4s, e4o, r2_guard;
    float fas, fao, fbs, fbo, uas, uao, ubs, ubo;   // PairTable fa, fb (scaled), ua, ub (list path)
    float rcore2s, rcore2o;                         // PairTable rcore2f (compensated form)
    float q2, qinv;                                 // 2^(2 qk), 2^-qk
};

__device__ inline PairF32 make_pair_f32(const PairTable &pt, int ci, const Grid &g) {
    PairF32 pc;
    pc.rc2s = pt.rc2f[ci] * g.q2f;
    pc.rc2o = pt.rc2f[2] * g.q2f;
    pc.sig2s = pt.sig2f[ci];
    pc.sig2o = pt.sig2f[2];
    pc.e24s = pt.eps24f[ci] * g.qinvf;
    pc.e24o = pt.eps24f[2] * g.qinvf;
    pc.e4s = pt.eps4f[ci];
    pc.e4o = pt.eps4f[2];
    pc.r2_guard = pt.r2_guard * g.q2f;
    pc.fas = pt.fa[ci] * g.qinvf;
    pc.fao = pt.fa[2] * g.qinvf;
    pc.fbs = pt.fb[ci] * g.qinvf;
    pc.fbo = pt.fb[2] * g.qinvf;
    pc.uas = pt.ua[ci];
    pc.uao = pt.ua[2];
    pc.ubs = pt.ub[ci];
    pc.ubo = pt.ub[2];
    pc.rcore2s = pt.rcore2f[ci] * g.q2f;
    pc.rcore2o = pt.rcore2f[2] * g.q2f;
    pc.q2 = g.q2f;
    pc.qinv = g.qinvf;
    return pc;
}

// Repulsive-core pairs of the compensated FP32 form.  A pair whose r^2 is
// below rcore2f (r < kCoreRatio sigma_ij) is left out of the FP32 sums and
// evaluated in FP64 from its exact fixed-point displacement.  Such pairs are
// deferred: the pass notes them (list rows in a row mask, cell candidates in
// a per-lane LDS queue of kCoreQ entries) and evaluates them after its FP32
// loop, in the order noted, so the FP32 loop stays branch-free and the FP64
// work runs once per noted pair instead of masked in every iteration.
constexpr int kCoreQ = 9;

// FP64 pair constants of one type pair (PairTable index).
struct CoreConst {
    double sig2, eps24, eps4, vc;
};
__device__ inline CoreConst core_const(const PairTable &pt, int pidx) {
    return CoreConst{pt.sig2[pidx], pt.eps24[pidx], pt.eps4[pidx], pt.vc[pidx]};
}
// The same for a per-lane class ci in {0, 1, 2}, by selects between the
// table's three (wave-uniform) entries rather than a per-lane load.
__device__ inline CoreConst core_const_sel(const PairTable &pt, int ci) {
    const CoreConst a = core_const(pt, 0), b = core_const(pt, 1), c = core_const(pt, 2);
    return CoreConst{ci == 0 ? a.sig2 : (ci == 1 ? b.sig2 : c.sig2), ci == 0 ? a.eps24 : (ci == 1 ? b.eps24 : c.eps24),
                     ci == 0 ? a.eps4 : (ci == 1 ? b.eps4 : c.eps4), ci == 0 ? a.vc : (ci == 1 ? b.vc : c.vc)};
}
__device__ inline CoreConst core_select(bool same, const CoreConst &a, const CoreConst &b) {
    return CoreConst{same ? a.sig2 : b.sig2, same ? a.eps24 : b.eps24, same ? a.eps4 : b.eps4,
                     same ? a.vc : b.vc};
}

// FP64 force (and energy sums) of one pair from its fixed-point displacement
// (dxi, dyi) * qinv, which converts exactly; added to (gx, gy).  1/r^2 from
// the hardware FP64 reciprocal refined by two Newton steps (relative error at
// the FP64 rounding level).  Only core pairs come here, well inside the
// cutoff.
template <int OBS>
__device__ inline void pair_core64(int dxi, int dyi, double qinv, const CoreConst &k, double &gx, double &gy,
                                   ObsSum &ob, bool want_obs) {
    const double dx = static_cast<double>(dxi) * qinv, dy = static_cast<double>(dyi) * qinv;
    const double r2 = __builtin_fma(dx, dx, dy * dy);
    if (!(r2 > 0.0)) return;   // coincident particles: no pair, as in the FP32 pass
    double y = __builtin_amdgcn_rcp(r2);
    y = __builtin_fma(__builtin_fma(-r2, y, 1.0), y, y);
    y = __builtin_fma(__builtin_fma(-r2, y, 1.0), y, y);
    const double sr2 = k.sig2 * y;
    const double sr6 = sr2 * sr2 * sr2;
    const double sr12 = sr6 * sr6;
    const double tmp = k.eps24 * __builtin_fma(2.0, sr12, -sr6) * y;
    const double fx = tmp * dx, fy = tmp * dy;
    gx += fx;
    gy += fy;
    if constexpr (OBS == SUMS_ENERGY) {
        if (want_obs) {
            ob.u += 0.5 * (k.eps4 * (sr12 - sr6) - k.vc);
            ob.w += 0.5 * (dx * fx + dy * fy);
            ob.lam += 0.5 * (fx * dx - fy * dy);
        }
    } else if constexpr (OBS == SUMS_VIRIAL) {
        if (want_obs) ob.w += 0.5 * (dx * fx + dy * fy);
    }
}

// Fixed-point displacement of the particle, whose image moved by minus the
// partner cell's offset is (pix, piy), from the partner image pj; on small
// grids (cells narrower than the cutoff, never in list mode) reduced to the
// minimum image along that axis.
__device__ inline void disp_fixed(int pix, int piy, int2 pj, const Grid &g, int &dxi, int &dyi) {
    dxi = pix - pj.x;
    dyi = piy - pj.y;
    if (g.small_x) dxi -= g.Lxq * static_cast<int>(rintf(static_cast<float>(dxi) / static_cast<float>(g.Lxq)));
    if (g.periodic_y && g.small_y)
        dyi -= g.Lyq * static_cast<int>(rintf(static_cast<float>(dyi) / static_cast<float>(g.Lyq)));
}

// FP32 pairs of a particle, whose fixed-point image moved by minus this
// neighbour cell's offset (ox, oy) is (pix, piy), with the candidates [k0, k1)
// of that cell, in ascending slot order.  The cell's pair forces are summed in
// FP32 and the cell total enters the particle's accumulator once: a
// compensated (two-sum) add in the COMP32 form, an FP64 add in the ACC64 form
// (which adds every pair in FP64, as the reference does).  Self is excluded by
// r^2 > 0.  UNI: all type pairs share sigma (rc and sigma^2 need no select);
// the arithmetic is identical to the general form.  COMP32: core pairs go to
// the caller's queue cq (one lane's column, entry j at cq[j * kBlock + lane]:
// slot k and the offset code (ox + 2) + 5 (oy + 2)) while it has room, else
// (and with LPP > 1) they are evaluated at once.  ACC64: pairs closer than
// r2_guard, whose FP32 (sigma/r)^12 could overflow, are evaluated in FP64.
template <int ACC, int OBS, bool UNI, int LPP>
__device__ __forceinline__ void pair_cell_f32(const BatchPtrs &P, uint32_t k0, uint32_t k1, int pix, int piy,
                                              int ci, int ox, int oy, const Grid &g, const PairTable &pt,
                                              const PairF32 &pc, ForceSum<ACC> &F, ObsSum &ob, bool want_obs,
                                              uint2 *cq, int &qn) {
    const int2 *__restrict__ img = static_cast<const int2 *>(P.img);
    const uint8_t *__restrict__ clsb = P.icls;
    float fxc = 0.0f, fyc = 0.0f;
    float wc = 0.0f;   // SUMS_VIRIAL, compensated form: the cell's sum of tmp r^2
    // with LPP lanes per particle, lane q of the group takes candidates
    // k0 + q, k0 + q + LPP, ...
    for (uint32_t k = k0 + (LPP > 1 ? threadIdx.x % LPP : 0u); k < k1; k += LPP) {
        const int2 pj = img[k];
        const int cj = clsb[k];
        int dxi, dyi;
        disp_fixed(pix, piy, pj, g, dxi, dyi);
        const float dx = static_cast<float>(dxi), dy = static_cast<float>(dyi);
        const float rr = __builtin_fmaf(dx, dx, dy * dy);
        const bool same = cj == ci;
        const float rc2 = UNI ? pc.rc2s : (same ? pc.rc2s : pc.rc2o);
        if (rr < rc2 && rr > 0.0f) {
            bool fp32 = true;
            if constexpr (ACC == ACC_COMP32) {
                const float rco = UNI ? pc.rcore2s : (same ? pc.rcore2s : pc.rcore2o);
                if (rr < rco) {
                    fp32 = false;
                    if (LPP == 1 && qn < kCoreQ) {
                        cq[qn * kBlock + threadIdx.x] = make_uint2(k, static_cast<uint32_t>((ox + 2) + 5 * (oy + 2)));
                        ++qn;
                    } else {   // several lanes per particle, or more than kCoreQ core partners
                        pair_core64<OBS>(dxi, dyi, g.qinv, core_const_sel(pt, pair_index(ci, cj)), F.gx, F.gy, ob,
                                         want_obs);
                    }
                }
            } else {
                if (rr < pc.r2_guard) {
                    fp32 = false;
                    double fx, fy;
                    pair_fp64<OBS>(static_cast<double>(dxi) * g.qinv, static_cast<double>(dyi) * g.qinv,
                                   pair_index(ci, cj), pt, fx, fy, ob, want_obs);
                    F.addd(fx, fy);
                }
            }
            if (fp32) {
                const float sig2 = UNI ? pc.sig2s : (same ? pc.sig2s : pc.sig2o);
                const float e24 = same ? pc.e24s : pc.e24o;
                const float r2i = fast_rcp(rr) * pc.q2;
                const float sr2 = sig2 * r2i;
                const float sr6 = sr2 * sr2 * sr2;
                const float sr12 = sr6 * sr6;
                const float tmp = e24 * __builtin_fmaf(2.0f, sr12, -sr6) * r2i;
                const float fx = tmp * dx, fy = tmp * dy;
                if constexpr (ACC == ACC_COMP32) {
                    fxc += fx;
                    fyc += fy;
                    if constexpr (OBS == SUMS_VIRIAL) wc = __builtin_fmaf(tmp, rr, wc);
                } else {
                    F.add(fx, fy);
                    if constexpr (OBS == SUMS_VIRIAL) {
                        if (want_obs) ob.w += 0.5 * g.qinv * static_cast<double>(tmp * rr);
                    }
                }
                if constexpr (OBS == SUMS_ENERGY) {
                    if (want_obs) {
                        const float e4 = same ? pc.e4s : pc.e4o;
                        const float u = e4 * (sr12 - sr6);
                        const double ddx = static_cast<double>(dx) * g.qinv, ddy = static_cast<double>(dy) * g.qinv;
                        const double dfx = fx, dfy = fy;
                        ob.u += 0.5 * (static_cast<double>(u) - pt.vc[pair_index(ci, cj)]);
                        ob.w += 0.5 * (ddx * dfx + ddy * dfy);
                        ob.lam += 0.5 * (dfx * ddx - dfy * ddy);
                    }
                }
            }
        }
    }
    if constexpr (ACC == ACC_COMP32) {
        F.add(fxc, fyc);
        if constexpr (OBS == SUMS_VIRIAL) {
            if (want_obs) ob.w += 0.5 * g.qinv * static_cast<double>(wc);
        }
    }
}

// The queued core pairs of a cell pass (pair_cell_f32), in queue order; pi is
// the particle's own fixed-point image.
template <int OBS>
__device__ inline void cell_core_queue(const BatchPtrs &P, int2 pi, int ci, const Grid &g, const PairTable &pt,
                                       const uint2 *cq, int qn, double &gx, double &gy, ObsSum &ob, bool want_obs) {
    const int2 *img = static_cast<const int2 *>(P.img);
    for (int j = 0; j < qn; ++j) {
        const uint2 e = cq[j * kBlock + threadIdx.x];
        const int ox = static_cast<int>(e.y % 5u) - 2, oy = static_cast<int>(e.y / 5u) - 2;
        int dxi, dyi;
        disp_fixed(pi.x - ox * g.wq, pi.y - oy * g.hq, img[e.x], g, dxi, dyi);
        pair_core64<OBS>(dxi, dyi, g.qinv, core_const_sel(pt, pair_index(ci, P.icls[e.x])), gx, gy, ob, want_obs);
    }
}

// One neighbour-list entry of a particle whose fixed-point image, moved by
// minus the entry's neighbour-cell offset, is (psx, psy), with the partner
// image q: the FP32 pair of pair_cell_f32 (the same displacement), its force
// evaluated in powers of 1/r^2 as F/r = r^-2 r^-6 (fa r^-6 - fb).  The force
// is computed unconditionally and kept only when the pair is inside the
// cutoff, so no branch is taken per entry.  CHK selects what else is checked:
//   0  nothing;
//   1  (ACC64) pairs closer than the FP64 guard are left out here; `flag`
//      collects, as a bit per lane, whether one of them is also farther than
//      0 (0 < r^2 < r2_guard, compared as IEEE bit patterns, which order like
//      the values for r^2 >= 0), and the caller evaluates those in FP64;
//   2  (COMP32) core pairs (r^2 < rcore2) are left out here; `flag` is set
//      to whether this entry is one.
// PURE: the wave's lists pair no two type classes (kListMixed clear), so
// every entry takes the lane's own-class constants.
template <int ACC, int OBS, bool UNI, int CHK, bool PURE = false>
__device__ __forceinline__ void list_pair_f32(uint32_t v, int2 q, int psx, int psy, int ci, const PairTable &pt,
                                              const PairF32 &pc, float &fxc, float &fyc, float &wc,
                                              ForceSum<ACC> &F, ObsSum &ob, bool want_obs, bool &flag) {
    const float dx = static_cast<float>(psx - q.x);
    const float dy = static_cast<float>(psy - q.y);
    const float rr = __builtin_fmaf(dx, dx, dy * dy);
    const bool same = PURE || (v & kListSame) != 0u;
    const float rc2 = UNI ? pc.rc2s : (same ? pc.rc2s : pc.rc2o);
    const float fa = same ? pc.fas : pc.fao;
    const float fb = same ? pc.fbs : pc.fbo;
    const float r2i = fast_rcp(rr) * pc.q2;
    const float r6i = r2i * r2i * r2i;
    const float t = __builtin_fmaf(fa, r6i, -fb) * r6i * r2i;
    bool in;
    if constexpr (CHK == 2) {
        const float rco = UNI ? pc.rcore2s : (same ? pc.rcore2s : pc.rcore2o);
        flag = rr < rco;
        in = rr < rc2 && !flag;
    } else if constexpr (CHK == 1) {
        in = rr < rc2 && rr >= pc.r2_guard;
        flag = flag || (__float_as_uint(rr) - 1u < __float_as_uint(pc.r2_guard) - 1u);
    } else {
        in = rr < rc2;
    }
    const float tmp = in ? t : 0.0f;
    if constexpr (ACC == ACC_COMP32) {   // the run's FP32 sum, one rounding per term
        fxc = __builtin_fmaf(tmp, dx, fxc);
        fyc = __builtin_fmaf(tmp, dy, fyc);
        if constexpr (OBS == SUMS_VIRIAL) wc = __builtin_fmaf(tmp, rr, wc);
    } else {
        F.add(tmp * dx, tmp * dy);
        if constexpr (OBS == SUMS_VIRIAL) {
            if (want_obs) ob.w += 0.5 * static_cast<double>(pc.qinv) * static_cast<double>(tmp * rr);
        }
    }
    if constexpr (OBS == SUMS_ENERGY) {
        if (want_obs && in) {
            const float fx = tmp * dx, fy = tmp * dy;
            const float uu = __builtin_fmaf(same ? pc.uas : pc.uao, r6i, -(same ? pc.ubs : pc.ubo)) * r6i;
            const double qi = pc.qinv;
            const double ddx = static_cast<double>(dx) * qi, ddy = static_cast<double>(dy) * qi;
            const double dfx = fx, dfy = fy;
            ob.u += 0.5 * (static_cast<double>(uu) - pt.vc[same ? ci : 2]);
            ob.w += 0.5 * (ddx * dfx + ddy * dfy);
            ob.lam += 0.5 * (dfx * ddx - dfy * ddy);
        }
    }
}

// The FP64 evaluation of the guarded pairs among U entries (0 < r^2 <
// r2_guard), in entry order; ps[u] is the particle's image moved by minus
// entry u's cell offset.
template <int U, int OBS>
__device__ __forceinline__ void list_pairs_guarded(const uint32_t (&v)[U], const int2 (&q)[U], const int2 (&ps)[U],
                                                   int ci, const PairTable &pt, const PairF32 &pc, double &gx,
                                                   double &gy, ObsSum &ob, bool want_obs) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int dxi = ps[u].x - q[u].x, dyi = ps[u].y - q[u].y;
        const float dx = static_cast<float>(dxi), dy = static_cast<float>(dyi);
        const float rr = __builtin_fmaf(dx, dx, dy * dy);
        if (rr < pc.r2_guard && rr > 0.0f) {
            const double qi = pc.qinv;
            double fx, fy;
            pair_fp64<OBS>(static_cast<double>(dxi) * qi, static_cast<double>(dyi) * qi,
                           (v[u] & kListSame) ? ci : 2, pt, fx, fy, ob, want_obs);
            gx += fx;
            gy += fy;
        }
    }
}

// Rows [e0, e1) of slot s's list with one lane per particle; the bounds are
// the same for the whole wave, so the loop is uniform.  Each pass takes four
// rows; the next pass's rows are loaded before this pass's pairs are
// evaluated (the last pass reloads its own rows).  Rows are addressed from a
// scalar base plus the lane's 32-bit byte offset.  shb: the workgroup's
// [9][kBlock] table of the particles' images moved by minus each
// neighbour-cell offset, read at the entry's code (list_shift_off).  CHK as
// list_pair_f32; with CHK == 2 (at most 32 rows) each row shifts its core flag
// into m from the right (m = 2 m + flag), so row e0 + j ends at bit
// e1 - e0 - 1 - j.  PURE as list_pair_f32.
template <int ACC, int OBS, bool UNI, int CHK, bool PURE>
__device__ __forceinline__ void list_rows_f32(const BatchPtrs &P, uint32_t s, uint32_t e0, uint32_t e1,
                                              const char *shb, int ci, const PairTable &pt, const PairF32 &pc,
                                              ForceSum<ACC> &F, ObsSum &ob, bool want_obs, float &fxc, float &fyc,
                                              float &wc, double &gx, double &gy, uint32_t &m) {
    constexpr int U = 4;
    const char *imgb = static_cast<const char *>(P.img);
    const char *lb = reinterpret_cast<const char *>(P.nlist);
    const size_t rowb = static_cast<size_t>(P.N) * 4u;   // bytes per row
    const uint32_t soff = s * 4u;
    const uint32_t lane8 = threadIdx.x * 8u;
    // a row's base address is the same for the whole wave: a raw buffer
    // descriptor over the row (base read back from the first lane, so it is
    // uniform to the compiler and lives in scalar registers; extent one row,
    // N * 4 < 2^29 bytes), loaded at the lane's 32-bit offset, so no per-lane
    // 64-bit address is formed.  Descriptor word 3 0x00020000: 32-bit data
    // format, the value gfx9-family untyped buffer loads take.
    auto row_load = [&](uint32_t e) {
        const uint64_t a = reinterpret_cast<uint64_t>(lb) + static_cast<uint64_t>(e) * rowb;
        const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(a));
        const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(a >> 32));
        char *row = reinterpret_cast<char *>((static_cast<uint64_t>(hi) << 32) | lo);
        const __amdgpu_buffer_rsrc_t rs =
            __builtin_amdgcn_make_buffer_rsrc(row, 0, static_cast<int>(rowb), 0x00020000);
        return static_cast<uint32_t>(__builtin_amdgcn_raw_buffer_load_b32(rs, static_cast<int>(soff), 0, 0));
    };
    uint32_t v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = row_load(e0 + u);
    for (uint32_t e = e0; e < e1; e += U) {
        int2 q[U], ps[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            q[u] = *reinterpret_cast<const int2 *>(imgb + (v[u] & kListOff));
            ps[u] = *reinterpret_cast<const int2 *>(shb + (list_shift_off(v[u]) | lane8));
        }
        const uint32_t en = e + U < e1 ? e + U : e;
        uint32_t vn[U];
#pragma unroll
        for (int u = 0; u < U; ++u) vn[u] = row_load(en + u);
        bool slow = false;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if constexpr (CHK == 2) {
                bool cf = false;
                list_pair_f32<ACC, OBS, UNI, CHK, PURE>(v[u], q[u], ps[u].x, ps[u].y, ci, pt, pc, fxc, fyc, wc, F,
                                                        ob, want_obs, cf);
                m = m + m + (cf ? 1u : 0u);   // one add with carry-in
            } else {
                list_pair_f32<ACC, OBS, UNI, CHK, PURE>(v[u], q[u], ps[u].x, ps[u].y, ci, pt, pc, fxc, fyc, wc, F,
                                                        ob, want_obs, slow);
            }
        }
        if constexpr (CHK == 1) {
            if (__any(slow)) list_pairs_guarded<U, OBS>(v, q, ps, ci, pt, pc, gx, gy, ob, want_obs);
        }
        if constexpr (ACC == ACC_COMP32) {
            if (e & U) {   // every eight rows
                F.add(fxc, fyc);
                fxc = fyc = 0.0f;
                if constexpr (OBS == SUMS_VIRIAL) {
                    ob.w += 0.5 * static_cast<double>(pc.qinv) * static_cast<double>(wc);
                    wc = 0.0f;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = vn[u];
    }
}

// Takes the first (lowest-numbered) row left in a core mask of list_rows_f32:
// mask m holds rows [base, base + cnt), row base + j at bit cnt - 1 - j.
__device__ inline bool core_take(uint32_t &m, uint32_t base, uint32_t cnt, uint32_t &e) {
    if (m == 0u) return false;
    const uint32_t pos = 31u - static_cast<uint32_t>(__builtin_clz(m));
    m ^= 1u << pos;
    e = base + cnt - 1u - pos;
    return true;
}

// The core rows of slot s's list marked by list_rows_f32 in m1 (rows [0, n1))
// and m2 (rows [32, 32 + n2)), in ascending row order, evaluated in FP64 from
// the exact fixed-point displacement (the shifted own image from the table
// shb, the partner's image reloaded).  A wave runs as many passes as its lane
// with the most core rows needs (usually none or one).
template <int OBS>
__device__ inline void list_core_rows(const BatchPtrs &P, uint32_t s, int ci, const Grid &g, const PairTable &pt,
                                      const char *shb, uint32_t m1, uint32_t n1, uint32_t m2, uint32_t n2,
                                      double &gx, double &gy, ObsSum &ob, bool want_obs) {
    if (!__any((m1 | m2) != 0u)) return;
    const CoreConst ks = core_const_sel(pt, ci), ko = core_const(pt, 2);
    const size_t stride = static_cast<size_t>(P.N);
    const char *imgb = static_cast<const char *>(P.img);
    const uint32_t lane8 = threadIdx.x * 8u;
    while (__any((m1 | m2) != 0u)) {
        uint32_t e = 0u;
        if (core_take(m1, 0u, n1, e) || core_take(m2, 32u, n2, e)) {
            const uint32_t v = P.nlist[static_cast<size_t>(e) * stride + s];
            const int2 q = *reinterpret_cast<const int2 *>(imgb + (v & kListOff));
            const int2 ps = *reinterpret_cast<const int2 *>(shb + (list_shift_off(v) | lane8));
            pair_core64<OBS>(ps.x - q.x, ps.y - q.y, g.qinv, core_select((v & kListSame) != 0u, ks, ko), gx, gy, ob,
                             want_obs);
        }
    }
}

// FP32 pairs of cell-ordered slot s from its neighbour list (list mode):
// entries in the cell order of the build, each carrying its neighbour cell's
// offset, so the displacement is formed exactly as in pair_cell_f32.  Pair
// forces are summed in FP32 over runs of eight entries, each run entering the
// compensated accumulator once (COMP32); ACC64 adds every pair in FP64.  With
// LPP lanes per particle, lane q takes rows q, q + LPP, ...; rows past the
// list stand for list_pad.  COMP32: core pairs are evaluated in FP64 (after
// the rows with one lane per particle, at once with several).  shift: the
// workgroup's [9][kBlock] LDS table (LPP == 1).
template <int ACC, int OBS, bool UNI, int LPP>
__device__ __forceinline__ void pair_list_f32(const BatchPtrs &P, uint32_t s, int2 pi, int ci, const Grid &g,
                                              const PairTable &pt, const PairF32 &pc, ForceSum<ACC> &F,
                                              ObsSum &ob, bool want_obs, int2 (*shift)[kBlock], uint32_t *stash) {
    constexpr int U = 4;
    float fxc = 0.0f, fyc = 0.0f;
    float wc = 0.0f;   // SUMS_VIRIAL, compensated form: the run's sum of tmp r^2
    double gx = 0.0, gy = 0.0;
    if constexpr (LPP == 1) {
        // the nine shifted images of each particle, one table row per cell
        // code (list_entry), the lane's own column
#pragma unroll
        for (int oy = -1; oy <= 1; ++oy)
#pragma unroll
            for (int ox = -1; ox <= 1; ++ox)
                shift[(ox + 1) + 3 * (oy + 1)][threadIdx.x] = make_int2(pi.x - ox * g.wq, pi.y - oy * g.hq);
        const char *shb = reinterpret_cast<const char *>(&shift[0][0]);
        const uint32_t nr = __builtin_amdgcn_readfirstlane(P.ncount[s]);
        const uint32_t n = nr & kListRows;
        const bool pure = (nr & kListMixed) == 0u;   // uniform
        uint32_t m1 = 0u, m2 = 0u;
        if constexpr (ACC == ACC_COMP32) {
            // core flags of rows [0, 32) in m1, of rows [32, n) in m2; m1 waits
            // in LDS while the rows past 32 (long lists only) run
            const uint32_t n1 = n < 32u ? n : 32u;
            if (pure)
                list_rows_f32<ACC, OBS, UNI, 2, true>(P, s, 0u, n1, shb, ci, pt, pc, F, ob, want_obs, fxc, fyc, wc,
                                                      gx, gy, m1);
            else
                list_rows_f32<ACC, OBS, UNI, 2, false>(P, s, 0u, n1, shb, ci, pt, pc, F, ob, want_obs, fxc, fyc, wc,
                                                       gx, gy, m1);
            if (n > 32u) {
                stash[threadIdx.x] = m1;
                list_rows_f32<ACC, OBS, UNI, 2, false>(P, s, 32u, n, shb, ci, pt, pc, F, ob, want_obs, fxc, fyc, wc,
                                                       gx, gy, m2);
                uint32_t t = threadIdx.x;
                __asm__ __volatile__("" : "+v"(t));   // a fresh LDS read, not m1 kept in a register
                m1 = stash[t];
            }
            list_core_rows<OBS>(P, s, ci, g, pt, shb, m1, n1, m2, n > 32u ? n - 32u : 0u, gx, gy, ob, want_obs);
        } else {
            if (nr & kListGuard)
                list_rows_f32<ACC, OBS, UNI, 1, false>(P, s, 0u, n, shb, ci, pt, pc, F, ob, want_obs, fxc, fyc, wc,
                                                       gx, gy, m1);
            else if (pure)
                list_rows_f32<ACC, OBS, UNI, 0, true>(P, s, 0u, n, shb, ci, pt, pc, F, ob, want_obs, fxc, fyc, wc,
                                                      gx, gy, m1);
            else
                list_rows_f32<ACC, OBS, UNI, 0, false>(P, s, 0u, n, shb, ci, pt, pc, F, ob, want_obs, fxc, fyc, wc,
                                                       gx, gy, m1);
        }
    } else {
        (void)shift;
        (void)stash;
        const char *imgb = static_cast<const char *>(P.img);
        const size_t stride = static_cast<size_t>(P.N);
        const uint32_t n = P.ncount[s] & kListRows;
        const uint32_t pad = list_pad(s);
        uint32_t run = 0;
        for (uint32_t e = threadIdx.x % LPP; e < n; e += U * LPP) {
            uint32_t v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t ee = e + u * LPP;
                v[u] = ee < n ? P.nlist[ee * stride + s] : pad;
            }
            int2 q[U], ps[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                q[u] = *reinterpret_cast<const int2 *>(imgb + (v[u] & kListOff));
                ps[u] = make_int2(pi.x - list_oxi(v[u]) * g.wq, pi.y - list_oyi(v[u]) * g.hq);
            }
            bool slow = false;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if constexpr (ACC == ACC_COMP32) {
                    bool core = false;
                    list_pair_f32<ACC, OBS, UNI, 2>(v[u], q[u], ps[u].x, ps[u].y, ci, pt, pc, fxc, fyc, wc, F,
                                                    ob, want_obs, core);
                    if (core)
                        pair_core64<OBS>(ps[u].x - q[u].x, ps[u].y - q[u].y, g.qinv,
                                         core_const(pt, (v[u] & kListSame) ? ci : 2), gx, gy, ob, want_obs);
                    if (++run == 8u) {
                        F.add(fxc, fyc);
                        fxc = fyc = 0.0f;
                        if constexpr (OBS == SUMS_VIRIAL) {
                            ob.w += 0.5 * g.qinv * static_cast<double>(wc);
                            wc = 0.0f;
                        }
                        run = 0;
                    }
                } else {
                    list_pair_f32<ACC, OBS, UNI, 1>(v[u], q[u], ps[u].x, ps[u].y, ci, pt, pc, fxc, fyc, wc, F,
                                                    ob, want_obs, slow);
                }
            }
            if constexpr (ACC != ACC_COMP32) {
                if (slow) list_pairs_guarded<U, OBS>(v, q, ps, ci, pt, pc, gx, gy, ob, want_obs);
            }
        }
    }
    if constexpr (ACC == ACC_COMP32) {
        F.add(fxc, fyc);
        if constexpr (OBS == SUMS_VIRIAL) ob.w += 0.5 * g.qinv * static_cast<double>(wc);
    }
    F.addd(gx, gy);
}

// FP64 pairs of particle pi (wrapped FP64 position) with the candidates
// [k0, k1) of one neighbour cell shifted by (shx, shy) = +-L across the
// periodic boundary; the oracle's operation order, so FP64 forces are
// bit-identical to the oracle's.
template <int OBS>
__device__ __forceinline__ void pair_cell_f64(const double2 *__restrict__ img, const uint8_t *__restrict__ cls,
                                              uint32_t k0, uint32_t k1, uint32_t self, const double2 &pi, int ci,
                                              double shx, double shy, const Grid &g, const PairTable &pt,
                                              ForceSum<ACC_FP64> &F, ObsSum &ob, bool want_obs) {
    for (uint32_t k = k0; k < k1; ++k) {
        if (k == self) continue;
        const double2 pj = img[k];
        const int cj = cls[k];
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
        if (pair_fp64<OBS>(dx, dy, pair_index(ci, cj), pt, fx, fy, ob, want_obs)) F.addd(fx, fy);
    }
}

// Sum of the LPP lanes' partial forces and observables of one particle
// (xor butterfly inside the wave; every lane of the group ends with the same
// totals: the two-sum error term is exact, the other additions commute).
template <int LPP, int ACC>
__device__ inline void combine_lanes(ForceSum<ACC> &F, ObsSum &ob) {
#pragma unroll
    for (int o = 1; o < LPP; o <<= 1) {
        if constexpr (ACC == ACC_COMP32) {
            const float sx = __shfl_xor(F.fx.s, o, 64), cx = __shfl_xor(F.fx.c, o, 64);
            const float sy = __shfl_xor(F.fy.s, o, 64), cy = __shfl_xor(F.fy.c, o, 64);
            F.fx.add(sx);
            F.fx.c += cx;
            F.fy.add(sy);
            F.fy.c += cy;
            F.gx += __shfl_xor(F.gx, o, 64);
            F.gy += __shfl_xor(F.gy, o, 64);
        } else {
            F.fx += __shfl_xor(F.fx, o, 64);
            F.fy += __shfl_xor(F.fy, o, 64);
        }
        ob.u += __shfl_xor(ob.u, o, 64);
        ob.w += __shfl_xor(ob.w, o, 64);
        ob.lam += __shfl_xor(ob.lam, o, 64);
    }
}

// Forces of every particle of every active slot (LPP lanes per cell-ordered
// slot, 3x3 neighbour cells in the oracle's order: rows -1, 0, +1, cells -1,
// 0, +1, ascending id), then the slot's mode: second kick, FIRE kick, walls,
// and its reduction.  OBS (PairSums) compiles the energy/virial/lambda sums
// or the virial alone, for the slots that take them (energy passes, NPH
// steps); UNI = true when every slot of the batch has one
// sigma for all type pairs.  A workgroup (kBlock * LPP threads) covers the
// same kBlock particles whatever LPP is; LPP > 1 only for the FP32 pair forms.
// LEAN: a rebuild window's lean step with step parity `lean_parity`, whose
// pair pass uses the list unless the slot has a particle beyond skin / 2
// (fallback[parity]) or no usable list, and then the 5x5 cells of the last
// cell list.  Other passes have parity 0.  (A compile-time LEAN keeps the
// other passes' code free of the lean branch: as a runtime argument it cost
// the cell pass about 13%.)
template <typename IT, int ACC, int OBS, bool UNI, int LPP, bool LEAN>
__global__ void __launch_bounds__(kBlock * LPP) k_pair(BatchPtrs P, int32_t lean_parity) {
    const int32_t parity = LEAN ? lean_parity : 0;
    using AT = std::conditional_t<ACC == ACC_COMP32, float2, double2>;
    constexpr bool kFp64 = std::is_same_v<IT, double2>;
    static_assert(!kFp64 || LPP == 1, "the FP64 form keeps the oracle's sequential order");
    static_assert(LPP == 1 || LPP == 2 || LPP == 4, "lanes per particle");
    const int gb = act_block(P, xcd_block(blockIdx.x, gridDim.x));
    const int r = P.blk2sys[gb];
    const SysParams &sp = P.sp[r];
    DevScalars &sc = P.sc[r];
    int pmode = sp.pmode;
    if (pmode == PM_FIRE && !sc.f_do_pair) pmode = PM_NONE;
    if (pmode == PM_NONE || sp.n <= 0) return;
    const bool want_obs = OBS != SUMS_NONE && (pmode == PM_OBS || pmode == PM_NPH);
    const Grid g = sc.g;
    const PairTable &pt = sp.pt;
    const int li = (gb - sp.boff) * kBlock + static_cast<int>(threadIdx.x) / LPP;
    const bool lead = LPP == 1 || threadIdx.x % LPP == 0;
    double r0 = 0.0, r1 = 0.0, r2v = 0.0;
    ForceSum<kFp64 ? ACC_FP64 : ACC> F;
    ObsSum ob;
    // the particle's storage index and class, loaded with its image so the
    // epilogue's velocity load does not wait on them
    uint32_t id = 0;
    int ci = 0;
    // one lane per particle, FP32 forms: the list pass's table of shifted
    // images, or the cell pass's queue of core pairs (kCoreQ uint2 per lane)
    __shared__ int2 lds_tab[(LPP == 1 && !kFp64) ? 9 : 1][kBlock];
    __shared__ uint32_t lds_stash[(LPP == 1 && !kFp64) ? kBlock : 1];
    static_assert(9 * sizeof(int2) >= kCoreQ * sizeof(uint2), "core queue fits the shift table");
    uint2 *cq = reinterpret_cast<uint2 *>(&lds_tab[0][0]);
    int qn = 0;
    if (li < sp.n) {
        const uint32_t s = P.cs[sp.coff] + static_cast<uint32_t>(li);
        id = P.sorted[s];
        const uint32_t lc = P.scell[s];
        const int cx = static_cast<int>(lc % static_cast<uint32_t>(g.ncx));
        const int cy = static_cast<int>(lc / static_cast<uint32_t>(g.ncx));
        ci = P.icls[s];
        const IT pi = static_cast<const IT *>(P.img)[s];
        PairF32 pc{};
        if constexpr (!kFp64) pc = make_pair_f32(pt, ci, g);
        bool by_list = false;
        if constexpr (!kFp64) {
            if (LEAN && (sc.fallback[parity] || !sc.list_on || !sc.list_built || sc.list_overflow)) {
                // lean step the list does not cover: 5x5 cells
                // (5 distinct cells per row and column, periodic; else the
                // forces are not exact and the run reports it)
                if ((g.ncx < 5 || g.ncy < 5 || !g.periodic_y) && li == 0) sc.violated = 1;
                for (int oy = -2; oy <= 2; ++oy) {
                    int yy = cy + oy;
                    if (yy < 0) yy += g.ncy;
                    else if (yy >= g.ncy) yy -= g.ncy;
                    const int rowbase = sp.coff + yy * g.ncx;
                    const int piy = pi.y - oy * g.hq;
                    for (int ox = -2; ox <= 2; ++ox) {
                        int xx = cx + ox;
                        if (xx < 0) xx += g.ncx;
                        else if (xx >= g.ncx) xx -= g.ncx;
                        const uint32_t k0 = P.cs[rowbase + xx], k1 = P.cs[rowbase + xx + 1];
                        pair_cell_f32<ACC, OBS, UNI, LPP>(P, k0, k1, pi.x - ox * g.wq, piy, ci, ox, oy, g, pt, pc, F,
                                                          ob, want_obs, cq, qn);
                    }
                }
                by_list = true;
            } else if (sc.list_on && !sc.list_overflow) {
                pair_list_f32<ACC, OBS, UNI, LPP>(P, s, pi, ci, g, pt, pc, F, ob, want_obs, lds_tab, lds_stash);
                by_list = true;
            }
        }
        for (int oy = -1; oy <= 1 && !by_list; ++oy) {
            int yy = cy + oy;
            double sy = 0.0;
            if (g.periodic_y) {
                if (yy < 0) {
                    yy += g.ncy;
                    sy = -g.Ly;
                } else if (yy >= g.ncy) {
                    yy -= g.ncy;
                    sy = g.Ly;
                }
            } else if (yy < 0 || yy >= g.ncy) {
                continue;
            }
            // the row's three cells (periodic in x), their ranges loaded together
            const int rowbase = sp.coff + yy * g.ncx;
            const int xl = cx == 0 ? g.ncx - 1 : cx - 1;
            const int xr = cx == g.ncx - 1 ? 0 : cx + 1;
            const uint32_t a0 = P.cs[rowbase + xl], a1 = P.cs[rowbase + xl + 1];
            const uint32_t b0 = P.cs[rowbase + cx], b1 = P.cs[rowbase + cx + 1];
            const uint32_t c0 = P.cs[rowbase + xr], c1 = P.cs[rowbase + xr + 1];
            if constexpr (kFp64) {
                const double2 *img = static_cast<const double2 *>(P.img);
                const double sxl = cx == 0 ? -g.Lx : 0.0;
                const double sxr = cx == g.ncx - 1 ? g.Lx : 0.0;
                pair_cell_f64<OBS>(img, P.icls, a0, a1, s, pi, ci, sxl, sy, g, pt, F, ob, want_obs);
                pair_cell_f64<OBS>(img, P.icls, b0, b1, s, pi, ci, 0.0, sy, g, pt, F, ob, want_obs);
                pair_cell_f64<OBS>(img, P.icls, c0, c1, s, pi, ci, sxr, sy, g, pt, F, ob, want_obs);
            } else {
                const int piy = pi.y - oy * g.hq;
                pair_cell_f32<ACC, OBS, UNI, LPP>(P, a0, a1, pi.x + g.wq, piy, ci, -1, oy, g, pt, pc, F, ob, want_obs,
                                                  cq, qn);
                pair_cell_f32<ACC, OBS, UNI, LPP>(P, b0, b1, pi.x, piy, ci, 0, oy, g, pt, pc, F, ob, want_obs, cq, qn);
                pair_cell_f32<ACC, OBS, UNI, LPP>(P, c0, c1, pi.x - g.wq, piy, ci, 1, oy, g, pt, pc, F, ob, want_obs,
                                                  cq, qn);
            }
        }
        if constexpr (!kFp64 && ACC == ACC_COMP32 && LPP == 1)
            cell_core_queue<OBS>(P, pi, ci, g, pt, cq, qn, F.gx, F.gy, ob, want_obs);
        if constexpr (LPP > 1) combine_lanes<LPP, kFp64 ? ACC_FP64 : ACC>(F, ob);
    }
    if (li < sp.n && lead) {
        const bool kicks = pmode == PM_NVE || pmode == PM_NPH || pmode == PM_PISTON || pmode == PM_NVT ||
                           pmode == PM_FIRE;
        double2 q = make_double2(0.0, 0.0);
        if (kicks) q = P.v[id];   // in flight while the acceleration is formed and stored
        double axd, ayd;
        if constexpr (kFp64 || ACC == ACC_FP64) {
            axd = F.X() / pt.m[ci];
            ayd = F.Y() / pt.m[ci];
        } else {
            axd = F.X() * pt.inv_m[ci];
            ayd = F.Y() * pt.inv_m[ci];
        }
        if (pmode == PM_WALLS || pmode == PM_PISTON) {   // src/md_cuda_common.cu:498-552
            const double y = P.x[id].y;
            const double k = sp.k_piston;
            double fw = 0.0, up = 0.0, e = 0.0;
            if (k > 0.0) {
                if (y > g.Ly) {
                    const double ov = y - g.Ly;
                    fw -= k * ov;
                    up += k * ov;
                    e += 0.5 * k * ov * ov;
                }
                if (y < 0.0) {
                    const double ov = -y;
                    fw += k * ov;
                    e += 0.5 * k * ov * ov;
                }
            }
            if (fw != 0.0) ayd += fw / pt.m[ci];
            r1 = up;
            r2v = e;
        }
        if (pmode != PM_OBS) {
            AT out;
            out.x = static_cast<decltype(out.x)>(axd);
            out.y = static_cast<decltype(out.y)>(ayd);
            static_cast<AT *>(P.a)[id] = out;
            // the kicks use the stored (possibly FP32-rounded) acceleration
            axd = static_cast<double>(out.x);
            ayd = static_cast<double>(out.y);
        }
        if (pmode == PM_NVE || pmode == PM_NPH || pmode == PM_PISTON) {
            q.x += 0.5 * sp.dt * axd;
            q.y += 0.5 * sp.dt * ayd;
            P.v[id] = q;
            if (pmode == PM_NPH) r0 = 0.5 * pt.m[ci] * (q.x * q.x + q.y * q.y);
        } else if (pmode == PM_NVT) {
            const double xi_mid = sc.xi_mid[parity];
            double vx = q.x, vy = q.y;
            vx += 0.5 * sp.dt * (axd - xi_mid * vx);
            vy += 0.5 * sp.dt * (ayd - xi_mid * vy);
            q.x = vx;
            q.y = vy;
            P.v[id] = q;
            r0 = 0.5 * pt.m[ci] * (vx * vx + vy * vy);
            r1 = pt.m[ci] * (vx * axd + vy * ayd);
            r2v = 0.5 * pt.m[ci] * (axd * axd + ayd * ayd);
        } else if (pmode == PM_FIRE) {
            const double fdt = sc.f_dt;
            q.x += fdt * axd;
            q.y += fdt * ayd;
            P.v[id] = q;
            r0 = sqrt(axd * axd + ayd * ayd);
            r1 = q.x * axd + q.y * ayd;
        }
        if (pmode == PM_NPH) {
            r1 = ob.w;
        } else if (pmode == PM_OBS) {
            r0 = ob.u;
            r1 = ob.w;
            r2v = ob.lam;
        }
    }
    const bool reduce = pmode == PM_OBS || pmode == PM_WALLS || pmode == PM_NVT || pmode == PM_NPH ||
                        pmode == PM_PISTON || pmode == PM_FIRE;
    if (reduce) {
        double vals[3] = {r0, r1, r2v};
        block_partials<3, kWaves * LPP>(vals, pmode == PM_FIRE ? 1u : 0u, P.partials, P.TB, gb);
        if (gb == sp.boff && threadIdx.x == 0) {
            sc.pair_pending = 1;
            sc.pair_parity = parity;
        }
    }
}


__global__ void k_clear_violations(BatchPtrs P) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < P.S) P.sc[r].violated = 0;
}

// --------------------------------------------------------------------------
// per-slot finalisation of the reductions
// --------------------------------------------------------------------------

enum FinStage : int32_t { FIN_PAIR = 1, FIN_KINETIC = 2, FIN_WALLS = 3 };

// One workgroup per slot: combines the slot's workgroup partials written by
// the preceding kernel and updates the slot's scalars.
__global__ void __launch_bounds__(kBlock) k_finalize(BatchPtrs P, int32_t stage, const int32_t *want) {
    const int r = blockIdx.x;
    const SysParams &sp = P.sp[r];
    DevScalars &sc = P.sc[r];
    if (sp.n <= 0 || sp.nblk <= 0) return;
    if (stage == FIN_PAIR) {
        if (sc.pair_pending) finalize_pair_pending(P, sp, sc);
    } else if (stage == FIN_KINETIC) {
        if (!want[r]) return;
        double out[1];
        slot_reduce<1>(0u, P.partials, P.TB, sp.boff, act_blocks(sp), out);
        if (threadIdx.x == 0) sc.K = out[0];
    } else if (stage == FIN_WALLS) {
        if (!want[r]) return;
        double out[2];
        slot_reduce<2>(0u, P.partials, P.TB, sp.boff, act_blocks(sp), out);
        if (threadIdx.x == 0) {
            sc.F_p = out[0];
            sc.wallE = out[1];
        }
    }
}

// --------------------------------------------------------------------------
// small kernels: kinetic energy, walls, FIRE set-up
// --------------------------------------------------------------------------

// K = sum 1/2 m v^2 per slot (src/md_cuda_common.cu:554-589); slots with
// want[r] == 0 are skipped.
__global__ void __launch_bounds__(kBlock) k_kinetic(BatchPtrs P, const int32_t *want) {
    const int gb = blockIdx.x;
    const int r = P.blk2sys[gb];
    const SysParams &sp = P.sp[r];
    if (!want[r] || sp.n <= 0) return;
    const int li = (gb - sp.boff) * kBlock + threadIdx.x;
    double k = 0.0;
    if (li < sp.n) {
        const uint32_t i = canon_index(P, sp, P.sc[r], li);
        const double2 q = P.v[i];
        k = 0.5 * sp.pt.m[P.cls[i]] * (q.x * q.x + q.y * q.y);
    }
    double vals[1] = {k};
    block_partials<1>(vals, 0u, P.partials, P.TB, gb);
}

// Piston wall observables without touching accelerations
// (src/md_env.cu:2148-2203 with update_particle_acc = false).
__global__ void __launch_bounds__(kBlock) k_walls(BatchPtrs P, const int32_t *want) {
    const int gb = blockIdx.x;
    const int r = P.blk2sys[gb];
    const SysParams &sp = P.sp[r];
    if (!want[r] || sp.n <= 0) return;
    const double Ly = P.sc[r].g.Ly, k = sp.k_piston;
    const int li = (gb - sp.boff) * kBlock + threadIdx.x;
    double up = 0.0, e = 0.0;
    if (li < sp.n && k > 0.0) {
        const double y = P.x[canon_index(P, sp, P.sc[r], li)].y;
        if (y > Ly) {
            const double ov = y - Ly;
            up += k * ov;
            e += 0.5 * k * ov * ov;
        }
        if (y < 0.0) {
            const double ov = -y;
            e += 0.5 * k * ov * ov;
        }
    }
    double vals[2] = {up, e};
    block_partials<2>(vals, 0u, P.partials, P.TB, gb);
}

// Pressure-tensor profile in nb slabs along y of slot r (the reference's
// local_pressure_tensor_profile_kernel, src/md_cuda_common.cu:879-1030, on one
// periodic domain): particle i adds m v_a v_b to the slab of its y; each
// ordered pair inside the cutoff adds 1/2 d_a F_b (d = r_i - r_j, F the FP64
// pair force on i) to the slab of the pair's midpoint y_i - d_y / 2 (wrapped
// when y is periodic, else clamped into [0, Ly)); slab = int(y / Ly * nb)
// clamped to [0, nb).  The pairs come from the slot's current cell list (3x3
// cells, displacements formed as in pair_cell_f64 from the wrapped FP64
// positions), so the FP64 force of every pair is the oracle's.  Sums are
// int64 fixed point, units of 2^-32 (each contribution rounded to the
// nearest unit, saturating at 2^62), so their value does not depend on the
// order of the atomics.  acc: [3][nb] (xx, yy, xy); LDS: [3][nb] when LDS_HIST.
constexpr double kProfScale = 4294967296.0;   // 2^32

__device__ inline unsigned long long prof_fixed(double v) {
    double q = rint(v * kProfScale);
    q = q > 4.6e18 ? 4.6e18 : (q < -4.6e18 ? -4.6e18 : q);
    return static_cast<unsigned long long>(static_cast<long long>(q));
}

__device__ inline int prof_bin(double y, double Ly, int nb, bool periodic) {
    if (periodic) {
        y = pbc_wrap(y, Ly);
    } else if (y < 0.0) {
        y = 0.0;
    } else if (y >= Ly && Ly > 0.0) {
        y = Ly * (1.0 - 1e-12);
    }
    const double q = Ly > 0.0 ? y / Ly : 0.0;
    int b = static_cast<int>(q * static_cast<double>(nb));
    return b < 0 ? 0 : (b >= nb ? nb - 1 : b);
}

template <bool LDS_HIST>
__global__ void __launch_bounds__(kBlock) k_pressure_profile(BatchPtrs P, int32_t r, int32_t nb,
                                                             unsigned long long *acc) {
    extern __shared__ unsigned long long phist[];
    const SysParams &sp = P.sp[r];
    const Grid g = P.sc[r].g;
    const PairTable &pt = sp.pt;
    const bool py = g.periodic_y != 0;
    unsigned long long *h = LDS_HIST ? phist : acc;
    if constexpr (LDS_HIST) {
        for (int j = threadIdx.x; j < 3 * nb; j += kBlock) phist[j] = 0ull;
        __syncthreads();
    }
    const int li = blockIdx.x * kBlock + threadIdx.x;
    if (li < sp.n) {
        const uint32_t s = P.cs[sp.coff] + static_cast<uint32_t>(li);
        const uint32_t id = P.sorted[s];
        const int ci = P.cls[id];
        double2 pi = P.x[id];
        pi.x = pbc_wrap(pi.x, g.Lx);
        if (py) pi.y = pbc_wrap(pi.y, g.Ly);
        const double2 vi = P.v[id];
        const double m = pt.m[ci];
        {
            const int b = prof_bin(pi.y, g.Ly, nb, py);
            atomicAdd(&h[b], prof_fixed(m * vi.x * vi.x));
            atomicAdd(&h[nb + b], prof_fixed(m * vi.y * vi.y));
            atomicAdd(&h[2 * nb + b], prof_fixed(m * vi.x * vi.y));
        }
        const uint32_t lc = P.scell[s];
        const int cx = static_cast<int>(lc % static_cast<uint32_t>(g.ncx));
        const int cy = static_cast<int>(lc / static_cast<uint32_t>(g.ncx));
        for (int oy = -1; oy <= 1; ++oy) {
            int yy = cy + oy;
            double sy = 0.0;
            if (py) {
                if (yy < 0) {
                    yy += g.ncy;
                    sy = -g.Ly;
                } else if (yy >= g.ncy) {
                    yy -= g.ncy;
                    sy = g.Ly;
                }
            } else if (yy < 0 || yy >= g.ncy) {
                continue;
            }
            for (int ox = -1; ox <= 1; ++ox) {
                int xx = cx + ox;
                double sx = 0.0;
                if (xx < 0) {
                    xx += g.ncx;
                    sx = -g.Lx;
                } else if (xx >= g.ncx) {
                    xx -= g.ncx;
                    sx = g.Lx;
                }
                const int c = sp.coff + yy * g.ncx + xx;
                for (uint32_t k = P.cs[c]; k < P.cs[c + 1]; ++k) {
                    if (k == s) continue;
                    const uint32_t jd = P.sorted[k];
                    double2 pj = P.x[jd];
                    pj.x = pbc_wrap(pj.x, g.Lx);
                    if (py) pj.y = pbc_wrap(pj.y, g.Ly);
                    double dx = (pi.x - pj.x) - sx;
                    double dy = (pi.y - pj.y) - sy;
                    if (g.small_x) {
                        const double d = pi.x - pj.x;
                        dx = d - g.Lx * round(d / g.Lx);
                    }
                    if (py && g.small_y) {
                        const double d = pi.y - pj.y;
                        dy = d - g.Ly * round(d / g.Ly);
                    }
                    double fx, fy;
                    ObsSum unused;
                    if (!pair_fp64<false>(dx, dy, pair_index(ci, P.cls[jd]), pt, fx, fy, unused, false)) continue;
                    const int b = prof_bin(pi.y - 0.5 * dy, g.Ly, nb, py);
                    atomicAdd(&h[b], prof_fixed(0.5 * dx * fx));
                    atomicAdd(&h[nb + b], prof_fixed(0.5 * dy * fy));
                    atomicAdd(&h[2 * nb + b], prof_fixed(0.5 * dx * fy));
                }
            }
        }
    }
    if constexpr (LDS_HIST) {
        __syncthreads();
        for (int j = threadIdx.x; j < 3 * nb; j += kBlock)
            if (phist[j] != 0ull) atomicAdd(&acc[j], phist[j]);
    }
}

// MDSimulation::middle_reflect_LG (src/md_env.cu:1457-1508, kernel
// src/md_cuda_common.cu:26-65) on slot r: y wrapped, x reflected into the
// central slab [xmin, xmin + width] with mirrored edges (period 2 width).
__global__ void __launch_bounds__(kBlock) k_middle_reflect(BatchPtrs P, int32_t r, double xmin, double width) {
    const SysParams &sp = P.sp[r];
    const int li = blockIdx.x * kBlock + threadIdx.x;
    if (li >= sp.n) return;
    const Grid g = P.sc[r].g;
    const int i = sp.off + li;
    double2 p = P.x[i];
    p.y = pbc_wrap(p.y, g.Ly);
    const double period = 2.0 * width;
    double ym = fmod(p.x - xmin, period);
    if (ym < 0.0) ym += period;
    if (ym > width) ym = period - ym;
    p.x = xmin + ym;
    P.x[i] = p;
}

// EnergyMinimizer::initialize_from_frame (src/md_energy_minimizer.cu:343-380):
// wrap positions, zero velocities and accelerations (the acceleration buffer
// is viewed as 16 bytes per particle whatever its precision).
__global__ void __launch_bounds__(kBlock) k_fire_init(BatchPtrs P, const int32_t *want) {
    const int gb = blockIdx.x;
    const int r = P.blk2sys[gb];
    const SysParams &sp = P.sp[r];
    if (!want[r]) return;
    const int li = (gb - sp.boff) * kBlock + threadIdx.x;
    if (li >= sp.n) return;
    const int i = sp.off + li;
    const Grid g = P.sc[r].g;
    double2 p = P.x[i];
    p.x = pbc_wrap(p.x, g.Lx);
    p.y = pbc_wrap(p.y, g.Ly);
    P.x[i] = p;
    P.v[i] = make_double2(0.0, 0.0);
    static_cast<double2 *>(P.a)[i] = make_double2(0.0, 0.0);
}

// --------------------------------------------------------------------------
// storage reorder (Batch::reorder) and the identity id maps
// --------------------------------------------------------------------------

// Slots with want[r]: dst[off + li] = src[sorted[off + li]], one per-particle
// array gathered into the slot's cell order (one thread per cell-ordered
// slot).  Every array is gathered with the same sorted[] before
// k_reorder_ids rewrites it.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_gather_cell_order(BatchPtrs P, const int32_t *want, const T *src,
                                                              T *dst) {
    const int gb = blockIdx.x;
    const int r = P.blk2sys[gb];
    const SysParams &sp = P.sp[r];
    if (!want[r]) return;
    const int li = (gb - sp.boff) * kBlock + threadIdx.x;
    if (li >= sp.n) return;
    const uint32_t k = static_cast<uint32_t>(sp.off + li);
    dst[k] = src[P.sorted[k]];
}

// Slots with want[r]: dst[k] = src[k] over their particles.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_copy_slots(BatchPtrs P, const int32_t *want, const T *src, T *dst) {
    const int gb = blockIdx.x;
    const int r = P.blk2sys[gb];
    const SysParams &sp = P.sp[r];
    if (!want[r]) return;
    const int li = (gb - sp.boff) * kBlock + threadIdx.x;
    if (li >= sp.n) return;
    const uint32_t k = static_cast<uint32_t>(sp.off + li);
    dst[k] = src[k];
}

// After the gathers: cell-ordered slot k holds the particle stored at k (the
// previous cell list, sorted_old, is the current one between steps), and the
// inverse map follows the gathered pid.
__global__ void __launch_bounds__(kBlock) k_reorder_ids(BatchPtrs P, const int32_t *want) {
    const int gb = blockIdx.x;
    const int r = P.blk2sys[gb];
    const SysParams &sp = P.sp[r];
    if (!want[r]) return;
    const int li = (gb - sp.boff) * kBlock + threadIdx.x;
    if (li >= sp.n) return;
    const uint32_t k = static_cast<uint32_t>(sp.off + li);
    P.sorted[k] = k;
    P.sorted_old[k] = k;
    P.sid[P.pid[k]] = k;
}

// pid[i] = sid[i] = i for i in [lo, hi): storage in load order.
__global__ void __launch_bounds__(kBlock) k_ident_ids(uint32_t *pid, uint32_t *sid, uint32_t lo, uint32_t hi) {
    const uint32_t i = lo + blockIdx.x * kBlock + threadIdx.x;
    if (i >= hi) return;
    pid[i] = i;
    sid[i] = i;
}

}  // namespace mdx
