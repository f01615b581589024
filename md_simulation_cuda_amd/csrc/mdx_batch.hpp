// Host-side batch of independent systems ("slots") sharing one set of device
// arrays, one HIP stream and one launch per kernel per step.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "mdx.h"
#include "mdx_sim.hpp"

namespace mdx {

int set_error(int code, const std::string &msg);

// Steps a slot runs between two reorders of its storage into cell order
// (Batch::reorder); reordering has no effect on results.
constexpr int kReorderSteps = 1000;

struct Slot {
    // capacity (fixed when the batch is built)
    int64_t n_cap = 0;
    int32_t ccap = 0;
    // occupancy
    bool loaded = false;
    int64_t n = 0;
    mdx_params p{};
    double rc_max = 0.0;
    std::vector<int32_t> type;
    std::vector<double> theta;
    double t = 0.0;
    double Lx0 = 0.0, Ly0 = 0.0;
    // box as the host last knew it; box_moved once an NPH or piston step may
    // have changed it on the device
    double Lx_now = 0.0, Ly_now = 0.0;
    bool box_moved = false;
    // validity of device-side quantities for the current positions
    bool cells_valid = false;
    int cells_py = 1;
    bool K_valid = false, W_valid = false;
    bool piston_active = false, nph_init = false, vp_init = false;
    int64_t since_reorder = 0;   // steps since the storage was last put in cell order
    // what the next step does
    int kind = SK_NONE;
    uint32_t flags = 0;
    double p_ext = 0.0;
};

// steps per graph launch when a call runs many steps of one mode
constexpr int kGraphSteps = 8;
// a windowed call of at least this many steps reads the violation flags at
// its end; shorter calls leave them for the next synchronising call
constexpr int kWindowCheckSteps = 64;
// most bins of an N profile (two int counters each in LDS)
constexpr int kProfileMaxBins = 8192;
// since_reorder of a slot whose storage is in frame order: due at once
constexpr int kReorderDue = 1 << 30;

struct Batch {
    int device = 0;
    hipStream_t stream = nullptr;       // the stream every launch and copy is issued on
    hipStream_t own_stream = nullptr;   // the stream the batch created (destroyed with it)
    int prec = MDX_PREC_COMP32;
    std::vector<Slot> slots;
    int64_t N = 0;     // total particle capacity
    int32_t C = 0;     // total cell capacity
    int32_t TB = 0;    // per-particle workgroups
    int32_t TCB = 0;   // per-cell workgroups
    int32_t TLB = 0;   // lean-step workgroups (kLeanSpan slots each)
    int32_t TA = 0;    // per-particle workgroups of loaded slots (BatchPtrs::actblk)
    int32_t TCA = 0;   // per-cell workgroups of loaded slots (BatchPtrs::actcblk)
    std::vector<SysParams> h_sp;
    std::vector<int32_t> h_want;
    // device memory
    double2 *x = nullptr, *v = nullptr;
    void *a = nullptr, *img = nullptr;
    uint8_t *cls = nullptr, *icls = nullptr;
    uint32_t *key = nullptr, *sorted = nullptr, *scell = nullptr, *count = nullptr, *cs = nullptr,
             *bsum = nullptr;
    uint32_t *sorted_old = nullptr, *cs_old = nullptr, *movers = nullptr, *mover_off = nullptr,
             *n_movers = nullptr;
    uint32_t *arrivals = nullptr;
    double *partials = nullptr;
    SysParams *d_sp = nullptr;
    DevScalars *d_sc = nullptr;
    int32_t *d_blk2sys = nullptr, *d_cblk2sys = nullptr, *d_lb2sys = nullptr, *d_want = nullptr;
    int32_t *d_actblk = nullptr, *d_actcblk = nullptr;
    std::vector<int32_t> h_actblk, h_actcblk;
    BatchPtrs ptrs{};
    // captured step graphs, by the pair pass's sums (PairSums: none, energy
    // sums, virial)
    hipGraphExec_t graph[3][4][2] = {};   // [PairSums][StepMode][step parity (lean steps)]
    hipGraphExec_t graph_window = nullptr;  // one whole rebuild window: the head step and window - 1 lean steps
    hipGraphExec_t graph_multi[3][4] = {};  // [PairSums][StepMode]: kGraphSteps steps of one mode (not lean)
    bool use_graph = true;
    bool graph_uni = true;   // pair-kernel variant the captured graphs hold
    int lpp = 1;             // lanes per particle of the FP32 pair kernel (1, 2, 4)
    double cell_scale = 1.0; // last cell_scale applied by set_tuning
    double skin = 0.0;       // neighbour-list skin (0: cell list every step)
    int list_cap = 64;       // neighbour-list entries per particle
    int window = 0;          // list mode: steps per rebuild window
    uint32_t *key_list = nullptr, *nlist = nullptr, *ncount = nullptr;
    uint32_t *pid = nullptr, *sid = nullptr;   // storage <-> original index maps
    int reorder_every = kReorderSteps;         // 0: storage stays in load order
    void *reorder_tmp = nullptr;              // 16 B per particle, allocated by the first reorder
    double2 *x0 = nullptr;
    int2 *img0 = nullptr;
    // rebuild windows persist across step calls: the step position within the
    // current window (0: the next step is a head step)
    int win_pos = 0;
    // the last pair pass's partials are not yet combined (run_windows leaves
    // them for k_box, k_lean or flush; the partials are consumed with the pair
    // mode they were produced under, so flush runs before SysParams change)
    bool fin_pending = false;
    // a windowed run ended without checking the slots' violation flags
    bool viol_unchecked = false;
    std::vector<SysParams> sp_dev;   // the SysParams the device holds
    // timing
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> force_ev, step_ev;
    double force_ms = 0.0, step_ms = 0.0;
    int64_t force_n = 0, step_n = 0;

    ~Batch();
    int init(int nslots, const int64_t *n_cap, const int32_t *ccap, int precision, int device_ordinal);

    // slot lifecycle
    int load(int r, const mdx_params &p, const mdx_particle64 *frame, int64_t n);
    int unload(int r);
    int set_state(int r, const mdx_particle64 *frame, int64_t n);

    // stepping: every loaded slot with kind != SK_NONE advances nsteps
    int step(int nsteps);
    int set_kind(int r, int kind, double p_ext, uint32_t flags);

    // passes at the current positions for the slots in `which`
    int compute_forces(const std::vector<int> &which, int periodic_y, bool walls);
    int observables(int r, uint32_t mask, mdx_obs *out);
    int partial_U_lambda(int r, double eps, double *out);
    int collect(int r, mdx_particle64 *out, int64_t cap);
    int cell_list(int r, int periodic_y, int32_t *ncx, int32_t *ncy, uint32_t *keys, uint32_t *cell_start,
                  uint32_t *sorted_ids, int64_t cap_n, int64_t cap_cells);
    int fire(int r, const mdx_fire_params &fp, double ftol, int max_steps, mdx_fire_result *res,
             int32_t *es, double *et, int cap);
    // FIRE as a per-slot kind of ordinary stepping: after fire_begin the slot
    // takes one FIRE iteration per step (the device-side schedule) until it
    // converges or reaches max_steps, alongside whatever the other slots do.
    int fire_begin(int r, const mdx_fire_params &fp, double ftol, int max_steps);
    // done = converged or max_steps iterations taken; fills res (PE only when done)
    int fire_poll(int r, int *done, mdx_fire_result *res);
    // K/U/P/virial of several slots from one kinetic and one energy pass
    int observables_many(const std::vector<int> &which, uint32_t mask, mdx_obs *out);
    // pair-pass tuning: cell size of every loaded slot and lanes per particle
    int set_tuning(const mdx_tuning &t);
    mdx_tuning tuning() const;
    // times the candidates on a copy of the batch (kind_override >= 0: every
    // loaded slot steps that kind) and applies the fastest
    int autotune(int steps, mdx_tune_report *rep, int kind_override);
    int set_precision(int precision);
    int read_scalars(int r, DevScalars &h);
    int sync();

    // internals
    int upload_sp(int r);
    int upload_all_sp();
    // obs: the pair pass's sums (PairSums)
    int launch_step_kernels(int obs, int mode, int parity);
    void launch_pair(int obs, int parity, bool lean, hipEvent_t e0, hipEvent_t e1);
    int time_pair_pass(int reps, double *avg_ms);   // pair kernel duration, launches back to back
    int set_stream(hipStream_t s);                  // issue on s (null: the batch's own stream)
    int run_passes(int count, int obs);
    int replay(int obs, int mode, int count, int parity = 0);
    int capture_steps(hipGraphExec_t &ge, int obs, int mode, int parity, int steps);
    int n_profile(int r, int nb, int32_t *out);
    int pressure_profile(int r, int nb, double *out);   // [Pxx | Pyy | Pxy] per slab along y
    int n_profile_noalloc(int r, int nb, int32_t *out);   // counters in the device's fixed scratch array
    int middle_reflect(int r);
    int run_windows(int count);
    bool windows_apply();
    int pre_step();
    int rebuild_cells(const std::vector<int> &which, int periodic_y);
    int kinetic(const std::vector<int> &which);
    int walls(const std::vector<int> &which);
    int obs_pass(int r, int periodic_y);
    void invalidate_graphs();
    int update_active();   // BatchPtrs::actblk and TA after a load or unload
    bool uniform_sigma() const;   // every loaded slot has one sigma for all type pairs
    // a slot of this ensemble kind steps in list mode: a positive skin, list
    // arrays, an FP32 form and NVE or NVT (skin 0 is the cell list every step,
    // also after an earlier tuning allocated the lists)
    bool list_kind(int kind) const;
    double cell_side_for(const Slot &s) const;   // cell side of a slot under the batch's tuning
    int ensure_list_arrays();
    int write_scalar_block(int r, const DevScalars &h);
    // slots r in `which`: storage put in the slot's cell order (no effect on results)
    int reorder(const std::vector<int> &which);
    int reorder_due();                        // the slots whose kReorderSteps have run
    int reset_ids(int r);                     // storage of slot r in load order
    int flush();                              // combine pending pair partials (k_finalize)
    int check_violations();                   // a lean step left exact forces' reach: MDX_ERR_STATE
};

}  // namespace mdx
