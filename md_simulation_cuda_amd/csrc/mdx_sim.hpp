// Per-system parameters (host-written) and device-resident scalars
// (device-written) of a batch of independent systems.
//
// A batch owns S slots.  Slot r owns particles [off_r, off_r + n_cap_r) of the
// concatenated particle arrays, cells [coff_r, coff_r + ccap_r] of the
// concatenated cell arrays, per-particle workgroups [boff_r, boff_r + nblk_r)
// and per-cell workgroups [cboff_r, cboff_r + ncblk_r).  Every kernel covers
// all slots in one launch; a workgroup belongs to exactly one slot, so each
// slot's reductions depend only on its own particles.  Cell coff_r + ccap_r is
// the slot's padding cell: its count is n_cap_r - n_r, so the cell starts of
// the batch-wide scan put slot r's cell-ordered particles at [off_r, off_r +
// n_r) whatever the other slots hold, and a slot that keeps its cell list
// between steps (list mode, idle) keeps valid cell starts while other slots
// are loaded or unloaded.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "mdx_device.hpp"

namespace mdx {

// Lean-step workgroups (rebuild windows): kLeanThreads threads covering
// kLeanSpan consecutive cell-ordered slots of one slot, kLeanPer per thread.
constexpr int kLeanThreads = 1024;
constexpr int kLeanPer = 4;
constexpr int kLeanSpan = kLeanThreads * kLeanPer;

enum StepKind : int32_t {
    SK_NONE = 0,     // slot idle this launch
    SK_NVE = 1,
    SK_NVT = 2,
    SK_NPH = 3,
    SK_PISTON = 4,
    SK_FIRE = 5,
    SK_KEYS = 6,     // rebuild the cell list at the current positions only
};

// Pair-kernel modes (per slot, runtime).
enum PairMode : int32_t {
    PM_NONE = 0,
    PM_PLAIN = 1,    // accelerations only (cal_forces)
    PM_WALLS = 2,    // accelerations + piston walls (cal_forces_piston)
    PM_OBS = 3,      // U, W, lambda sums; no writes to particles
    PM_NVE = 4,
    PM_NVT = 5,
    PM_NPH = 6,
    PM_PISTON = 7,
    PM_FIRE = 8,
};

struct SysParams {
    int32_t n;           // particles in the slot (<= n_cap)
    int32_t off;         // first particle
    int32_t n_cap;
    int32_t coff;        // first cell
    int32_t ccap;        // cell capacity
    int32_t boff, nblk;  // per-particle workgroups
    int32_t cboff, ncblk;
    // the cells the slot's grid may use while its job is loaded (at most
    // ccap): the cell-side step kernels cover these, and k_box keeps the grid
    // within them
    int32_t cbound;
    int32_t lboff, nlblk;  // lean-step workgroups (kLeanSpan cell-ordered slots each)
    int32_t kind;        // StepKind of the drift kernel for this slot
    int32_t pmode;       // PairMode of the pair kernel for this slot
    int32_t periodic_y;  // grid of the step: 1 except piston
    int32_t middle_wrap;
    int32_t wrap_keys;   // SK_KEYS: wrap positions for keys/image
    double cell_side;    // cell side target = cell_scale * rc_max
    double rc_max;
    double dt, T0, Q, g_dof;
    double slab_xmin, slab_period, slab_width;   // middle_reflect_LG
    double p_ext, W_baro, k_piston;
    // FIRE
    double ftol, f_inc, f_dec, dt_max, alpha0, max_step;
    int32_t N_min, f_max_steps;
    // neighbour-list mode (FP32 forms, NVE/NVT): each particle lists the
    // partners within r_list = rc_max + skin; the list (and the cell list it is
    // built from) is rebuilt at the step where some particle of the slot has
    // moved more than skin / 2 since the last build
    int32_t list_mode, list_cap;
    double skin_half2;   // (skin / 2)^2
    double r_list;       // rc_max + skin
    PairTable pt;
};

// Device-written scalars of one slot.
struct DevScalars {
    Grid g;              // current cell grid (box lives in g.Lx, g.Ly)
    Grid g_old;          // grid of the previous cell list
    int32_t list_valid;  // the previous cell list (cs_old, sorted_old) holds every particle
    int32_t pair_pending;  // the last pair pass left partials that k_box / k_finalize / k_lean combine
    int32_t pair_parity;   // the step parity that pass ran with (0 outside rebuild windows)
    int32_t pad_s;
    double xi;           // Nose-Hoover variable after the last completed half update
    double eta;          // integral of xi dt: sum over NVT steps of dt (xi1 + xi_mid) / 2
    // xi of a step in flight (after both half updates), by step parity: steps
    // outside rebuild windows use [0]; a window's lean steps alternate, each
    // reading the previous step's entry while writing its own
    double xi_mid[2];
    double K;            // kinetic energy of the current velocities
    double K_half;       // kinetic energy after the first kick
    // sums of the last Nose-Hoover pair pass: sum m v.a and sum m |a|^2 / 2 at
    // the stored accelerations, from which a lean step gets K_half in closed form
    double S_va, S_aa;
    double W, U, lam, wallE;
    double p_A, gamma, nph_lin, nph_vsc, P_old;
    double v_p, F_p, last_P;
    double Ly_floor, Ly_ceil;
    int32_t clamped, box_error;
    // FIRE schedule (src/md_energy_minimizer.cu:170-203)
    double f_dt, f_alpha, f_maxF, f_power;
    int32_t f_streak, f_done, f_zero_v, f_step;
    int32_t f_conv_step, f_drifted, f_do_drift, f_do_pair;
    // neighbour-list mode
    int32_t list_on;        // this step uses the neighbour list
    int32_t rebuild;        // this step rebuilds the cell list (always outside list mode)
    int32_t list_built;     // the cell list was built from key_list (a list-mode build); nlist and x0 match it
    int32_t list_overflow;  // a particle had more than list_cap partners: list mode off
    // NPH in list mode: the factor the barostat has scaled positions by since
    // the last list build (the product of the steps' nph_lin)
    double lin_cum;
    // by step parity: a lean step with a particle beyond skin / 2 (set by
    // k_lean, cleared by the step before); its pair pass takes the 5x5 cells of
    // the last cell list instead of the list
    int32_t fallback[2];
    // a lean step with a particle beyond the reach of that 5x5 pass
    // ((2 w - rc_max) / 2), or a grid too small for it: its forces are not exact
    int32_t violated;
    // an unloaded slot whose cells a step has cleared (zero when loaded)
    int32_t cleared;
};

// The per-particle workgroups of a slot that hold its particles, the ones
// the step kernels run (BatchPtrs::actblk) and whose reduction partials the
// slot's sums combine.
__host__ __device__ inline int act_blocks(const SysParams &sp) { return sp.n > 0 ? (sp.n + kBlock - 1) / kBlock : 0; }

// Pointers of a batch, passed by value to every kernel.
struct BatchPtrs {
    double2 *x, *v;
    void *a;
    uint8_t *cls;
    uint32_t *key, *count, *cs, *sorted, *scell;
    void *img;
    uint8_t *icls;
    double *partials;       // [3][TB]
    uint32_t *sorted_old, *cs_old;   // previous cell list
    // per drift workgroup gb: movers[gb * kBlock + j], j < n_movers[gb], are
    // the particles whose cell changed; mover_off their arrival index
    uint32_t *movers, *mover_off;
    uint32_t *n_movers;
    uint32_t *arrivals;              // per cell: movers that arrived this step (zero at rest)
    // neighbour-list mode: key and position of each particle at the last
    // build; per cell-ordered slot s, entry e of its list at nlist[e * N + s]
    // (list_entry in mdx_kernels.hpp: partner slot, neighbour-cell offset,
    // same-class bit) and its row count ncount[s]
    uint32_t *key_list;
    double2 *x0;
    int2 *img0;       // per cell-ordered slot: its fixed-point image at the last list-mode build
    uint32_t *nlist, *ncount;
    // storage order: particle data (x, v, a, cls, key, key_list, x0) sits at
    // storage index i; pid[i] is its original (load-order) global index and
    // sid[pid[i]] = i.  The cell order sorts by pid, so it, and every result,
    // is independent of the storage order, which a slot's periodic reorder
    // (Batch::reorder) sets to its cell order for coalesced access.
    uint32_t *pid, *sid;
    const SysParams *sp;
    DevScalars *sc;
    const int32_t *blk2sys;   // per-particle workgroup -> slot
    // the per-particle workgroups that hold particles of loaded slots
    // (ceil(n / kBlock) from each slot's first), in order: the step kernels
    // launch one workgroup per entry and take their workgroup id from it
    // (act_block); act_all: every workgroup is in it, in order, so the id is
    // the launch index itself and the list is not read
    const int32_t *actblk;
    int32_t act_all;
    // the same for the per-cell workgroups: those of loaded slots up to
    // ceil(cbound / kBlock) from each slot's first
    const int32_t *actcblk;
    int32_t act_call;
    const int32_t *cblk2sys;  // per-cell workgroup -> slot
    const int32_t *lb2sys;    // lean-step workgroup -> slot
    int32_t TB;               // total per-particle workgroups
    int32_t S;                // slots
    int32_t C;                // total cells
    int32_t N;                // total particle capacity (stride of nlist)
};

// Workgroup id of launch index i of a step kernel launched over the active
// workgroups (BatchPtrs::actblk).
__device__ inline int act_block(const BatchPtrs &P, int i) { return P.act_all ? i : P.actblk[i]; }
__device__ inline int act_cblock(const BatchPtrs &P, int i) { return P.act_call ? i : P.actcblk[i]; }

// Canonical grid of a box (the oracle evaluates the same expressions).
__host__ __device__ inline int grid_dim(double L, double side, bool periodic) {
    double q = floor(L / side);
    const int lo = periodic ? 3 : 1;
    if (!(q >= lo)) return lo;
    if (q > 1.0e8) q = 1.0e8;
    return static_cast<int>(q);
}

__host__ __device__ inline Grid make_grid(double Lx, double Ly, double side, double rc_max, int periodic_y) {
    Grid g{};
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
    g.small_x = g.w < rc_max ? 1 : 0;
    g.small_y = g.h < rc_max ? 1 : 0;
    int e = 0;
    frexp(5.0 * (g.w > g.h ? g.w : g.h), &e);   // 5 max(w, h) < 2^e
    g.qk = 31 - e;
    g.qs = ldexp(1.0, g.qk);
    g.qinv = ldexp(1.0, -g.qk);
    g.q2f = static_cast<float>(ldexp(1.0, 2 * g.qk));
    g.qinvf = static_cast<float>(g.qinv);
    g.wq = to_fixed(g.w, g.qs);
    g.hq = to_fixed(g.h, g.qs);
    g.Lxq = to_fixed(Lx, g.qs);
    g.Lyq = to_fixed(Ly, g.qs);
    return g;
}

}  // namespace mdx
