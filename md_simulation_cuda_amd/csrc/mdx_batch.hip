// Batch of independent LJ systems on one GPU: allocation, slot admission,
// launch sequences per ensemble, observables, FIRE, hipGraph capture.
//
// Step sequencing restates src/md_env.cu:1576-2023 (NVE, Nose-Hoover, NPH,
// piston) and src/md_energy_minimizer.cu:122-272 (FIRE) for one periodic
// domain per slot: the MPI slab exchange of the reference (src/md_env.cu:
// 757-1274) is replaced by a cell list rebuilt on the device every step.
#include <hip/hip_ext.h>

#include "mdx_batch.hpp"
#include "mdx_kernels.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

namespace mdx {

namespace {
thread_local std::string g_err;
}

int set_error(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

const char *last_error() { return g_err.c_str(); }

#define MDX_HIP(...)                                                                                \
    do {                                                                                            \
        hipError_t e_ = (__VA_ARGS__);                                                              \
        if (e_ != hipSuccess)                                                                       \
            return set_error(MDX_ERR_HIP, std::string(#__VA_ARGS__) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define MDX_TRY(...)                   \
    do {                               \
        int rc_ = (__VA_ARGS__);       \
        if (rc_ != MDX_OK) return rc_; \
    } while (0)

namespace {

int ceil_div(int64_t a, int64_t b) { return static_cast<int>((a + b - 1) / b); }

PairTable make_pair_table(const mdx_params &p, double &rc_max) {
    PairTable pt{};
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
        // FP32 constants formed as src/md_cuda_common.cu:220-241 forms them
        const float sf = static_cast<float>(sig[k]), ef = static_cast<float>(eps[k]);
        const float rcf = cutf * sf;
        pt.rc2f[k] = rcf * rcf;
        pt.sig2f[k] = sf * sf;
        pt.eps24f[k] = 24.0f * ef;
        pt.eps4f[k] = 4.0f * ef;
        const double q6 = pt.sig2[k] * pt.sig2[k] * pt.sig2[k];   // sigma^6
        pt.fa[k] = static_cast<float>(2.0 * pt.eps24[k] * q6 * q6);
        pt.fb[k] = static_cast<float>(pt.eps24[k] * q6);
        pt.ua[k] = static_cast<float>(pt.eps4[k] * q6 * q6);
        pt.ub[k] = static_cast<float>(pt.eps4[k] * q6);
        const double rco = kCoreRatio * sig[k];
        pt.rcore2f[k] = static_cast<float>(rco * rco);
    }
    pt.m[0] = p.mass_a;
    pt.m[1] = p.mass_b;
    pt.m[2] = p.mass_b;
    for (int k = 0; k < 3; ++k) pt.inv_m[k] = 1.0 / pt.m[k];
    // FP32 (sigma/r)^12 / r^2 overflows below r ~ 2.3e-3 sigma: such
    // overlapping pairs are evaluated in FP64.
    const double smin = std::min(sig[0], std::min(sig[1], sig[2]));
    pt.r2_guard = static_cast<float>(1.0e-4 * smin * smin);
    rc_max = p.cutoff * std::max(sig[0], std::max(sig[1], sig[2]));
    return pt;
}

int validate_params(const mdx_params &p) {
    auto pos = [](double v) { return v > 0.0 && std::isfinite(v); };
    if (!pos(p.Lx) || !pos(p.Ly)) return set_error(MDX_ERR_BOX, "box sides must be positive and finite");
    if (!pos(p.sigma_aa) || !pos(p.sigma_bb) || !pos(p.sigma_ab))
        return set_error(MDX_ERR_INVALID, "sigma must be > 0");
    if (!pos(p.mass_a) || !pos(p.mass_b)) return set_error(MDX_ERR_INVALID, "masses must be > 0");
    if (!pos(p.cutoff)) return set_error(MDX_ERR_INVALID, "cutoff must be > 0");
    if (!std::isfinite(p.dt)) return set_error(MDX_ERR_INVALID, "dt must be finite");
    return MDX_OK;
}

}  // namespace

int mdx_cell_capacity(const mdx_params &p, double ratio) {
    double rc_max = 0.0;
    (void)make_pair_table(p, rc_max);
    const double side = std::max(p.cell_scale, 1.0) * rc_max;
    const double f = std::max(1.0, ratio);
    const Grid g = make_grid(p.Lx * f, p.Ly * f, side, rc_max, 1);
    const int64_t c = static_cast<int64_t>(g.ncx) * g.ncy;
    return c > (int64_t{1} << 30) ? -1 : static_cast<int>(c);
}

Batch::~Batch() {
    if (device >= 0) (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    invalidate_graphs();
    void *ptrs_[] = {x, v, a, img, cls, icls, key, sorted, scell, count, cs, bsum, partials,
                     d_sp, d_sc, d_blk2sys, d_cblk2sys, d_lb2sys, d_want, d_actblk, d_actcblk, sorted_old, cs_old,
                     movers, mover_off,
                     n_movers, arrivals, key_list, x0, img0, nlist, ncount, pid, sid, reorder_tmp};
    for (void *p : ptrs_)
        if (p) (void)hipFree(p);
    for (auto *vec : {&force_ev, &step_ev})
        for (auto &e : *vec) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
    if (own_stream && own_stream != stream) (void)hipStreamSynchronize(own_stream);
    if (own_stream) (void)hipStreamDestroy(own_stream);
}

// Issues every later launch and copy on s, a stream of the batch's device
// owned by the caller (null: back to the batch's own stream).  The work
// already issued completes first; captured graphs do not depend on the
// stream they are launched on and stay valid.
int Batch::set_stream(hipStream_t s) {
    MDX_HIP(hipSetDevice(device));
    if (s) {
        int dev = -1;
        MDX_HIP(hipStreamGetDevice(s, &dev));
        if (dev != device) return set_error(MDX_ERR_INVALID, "stream of another device");
    }
    MDX_TRY(flush());
    MDX_HIP(hipStreamSynchronize(stream));
    stream = s ? s : own_stream;
    return MDX_OK;
}

void Batch::invalidate_graphs() {
    for (auto &row : graph)
        for (auto &pair : row)
            for (auto &g : pair)
                if (g) {
                    (void)hipGraphExecDestroy(g);
                    g = nullptr;
                }
    if (graph_window) {
        (void)hipGraphExecDestroy(graph_window);
        graph_window = nullptr;
    }
    for (auto &row : graph_multi)
        for (auto &g : row)
            if (g) {
                (void)hipGraphExecDestroy(g);
                g = nullptr;
            }
}

// The per-particle workgroups of loaded slots (TA of them), for the step
// kernels' launches: free slots and the unfilled end of a slot take no
// workgroups.  Captured graphs hold the launch size, so a change of TA
// invalidates them; the list itself is read at run time.
int Batch::update_active() {
    h_actblk.clear();
    h_actcblk.clear();
    for (size_t r = 0; r < slots.size(); ++r) {
        if (!slots[r].loaded) continue;
        const int nb = ceil_div(slots[r].n, kBlock);
        for (int b = 0; b < nb; ++b) h_actblk.push_back(h_sp[r].boff + b);
        const int ncb = ceil_div(h_sp[r].cbound, kBlock);
        for (int b = 0; b < ncb; ++b) h_actcblk.push_back(h_sp[r].cboff + b);
    }
    const int32_t ta = static_cast<int32_t>(h_actblk.size()), tca = static_cast<int32_t>(h_actcblk.size());
    // ordered after the work in flight on the stream, which may still read the old lists
    if (ta > 0)
        MDX_HIP(hipMemcpyAsync(d_actblk, h_actblk.data(), h_actblk.size() * 4, hipMemcpyHostToDevice, stream));
    if (tca > 0)
        MDX_HIP(hipMemcpyAsync(d_actcblk, h_actcblk.data(), h_actcblk.size() * 4, hipMemcpyHostToDevice, stream));
    MDX_HIP(hipStreamSynchronize(stream));
    if (ta != TA || tca != TCA) invalidate_graphs();
    TA = ta;
    TCA = tca;
    ptrs.act_all = TA == TB ? 1 : 0;     // then the list is 0, 1, ..., TB - 1
    ptrs.act_call = TCA == TCB ? 1 : 0;
    return MDX_OK;
}

int Batch::init(int nslots, const int64_t *n_cap, const int32_t *ccap, int precision, int device_ordinal) {
    if (nslots <= 0 || nslots > (1 << 20)) return set_error(MDX_ERR_INVALID, "slot count out of range");
    if (precision < MDX_PREC_COMP32 || precision > MDX_PREC_FP64)
        return set_error(MDX_ERR_INVALID, "unknown precision");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return set_error(MDX_ERR_NODEVICE, "no HIP device");
    if (device_ordinal >= ndev) return set_error(MDX_ERR_INVALID, "device ordinal out of range");
    if (device_ordinal >= 0) MDX_HIP(hipSetDevice(device_ordinal));
    MDX_HIP(hipGetDevice(&device));
    MDX_HIP(hipStreamCreateWithFlags(&own_stream, hipStreamNonBlocking));
    stream = own_stream;
    prec = precision;
    slots.assign(nslots, Slot{});
    h_sp.assign(nslots, SysParams{});
    h_want.assign(nslots, 0);
    std::vector<int32_t> b2s, cb2s, lb2s;
    int64_t off = 0, coff = 0;
    int boff = 0, cboff = 0, lboff = 0;
    for (int r = 0; r < nslots; ++r) {
        if (n_cap[r] < 0 || ccap[r] < 0) return set_error(MDX_ERR_INVALID, "negative capacity");
        Slot &s = slots[r];
        s.n_cap = n_cap[r];
        // the slot's cells and its padding cell fill whole scan chunks, so
        // that the cell-start scan can skip the slots that do not rebuild
        const int64_t cc = (static_cast<int64_t>(ccap[r]) + kScanChunk) / kScanChunk * kScanChunk - 1;
        if (cc > INT32_MAX / 2) return set_error(MDX_ERR_CAPACITY, "batch above 2^29 particles or 2^30 cells");
        s.ccap = static_cast<int32_t>(cc);
        SysParams &sp = h_sp[r];
        sp.off = static_cast<int32_t>(off);
        sp.n_cap = static_cast<int32_t>(n_cap[r]);
        sp.coff = static_cast<int32_t>(coff);
        sp.ccap = s.ccap;
        sp.boff = boff;
        sp.nblk = ceil_div(n_cap[r], kBlock);
        sp.cboff = cboff;
        sp.ncblk = ceil_div(s.ccap, kBlock);   // (s.ccap + 1) / kBlock: cell block g holds cells kBlock g on
        sp.lboff = lboff;
        sp.nlblk = ceil_div(n_cap[r], kLeanSpan);
        sp.periodic_y = 1;
        sp.wrap_keys = 1;
        for (int b = 0; b < sp.nblk; ++b) b2s.push_back(r);
        for (int b = 0; b < sp.ncblk; ++b) cb2s.push_back(r);
        for (int b = 0; b < sp.nlblk; ++b) lb2s.push_back(r);
        off += n_cap[r];
        coff += s.ccap + 1;   // + the slot's padding cell (mdx_sim.hpp)
        boff += sp.nblk;
        cboff += sp.ncblk;
        lboff += sp.nlblk;
        // FP32 images are addressed with 32-bit byte offsets (slot << 3)
        if (off > (int64_t{1} << 29) - kBlock || coff > INT32_MAX / 2)
            return set_error(MDX_ERR_CAPACITY, "batch above 2^29 particles or 2^30 cells");
    }
    N = off;
    C = static_cast<int32_t>(coff);
    TB = boff;
    TCB = cboff;
    TLB = lboff;
    const size_t n16 = static_cast<size_t>(std::max<int64_t>(N, 1)) * 16;
    const size_t n4 = static_cast<size_t>(std::max<int64_t>(N, 1)) * 4;
    const size_t n1 = static_cast<size_t>(std::max<int64_t>(N, 1));
    MDX_HIP(hipMalloc(&x, n16));
    MDX_HIP(hipMalloc(&v, n16));
    MDX_HIP(hipMalloc(&a, n16));
    MDX_HIP(hipMalloc(&img, n16));
    MDX_HIP(hipMalloc(&cls, n1));
    MDX_HIP(hipMalloc(&icls, n1));
    MDX_HIP(hipMalloc(&key, n4));
    MDX_HIP(hipMalloc(&sorted, n4));
    MDX_HIP(hipMalloc(&scell, n4));
    MDX_HIP(hipMalloc(&count, (static_cast<size_t>(C) + 1) * 4));
    MDX_HIP(hipMalloc(&sorted_old, n4));
    MDX_HIP(hipMalloc(&cs_old, (static_cast<size_t>(C) + 1) * 4));
    // per drift workgroup: up to kBlock movers and their count
    const size_t mcap = static_cast<size_t>(std::max(TB, 1)) * kBlock * 4;
    MDX_HIP(hipMalloc(&movers, mcap));
    MDX_HIP(hipMalloc(&mover_off, mcap));
    MDX_HIP(hipMalloc(&n_movers, static_cast<size_t>(std::max(TB, 1)) * 4));
    MDX_HIP(hipMemset(n_movers, 0, static_cast<size_t>(std::max(TB, 1)) * 4));
    MDX_HIP(hipMalloc(&arrivals, std::max<size_t>(static_cast<size_t>(C), 1) * 4));
    MDX_HIP(hipMemset(arrivals, 0, std::max<size_t>(static_cast<size_t>(C), 1) * 4));
    MDX_HIP(hipMemset(key, 0, n4));
    MDX_HIP(hipMalloc(&cs, (static_cast<size_t>(C) + 1) * 4));
    MDX_HIP(hipMalloc(&bsum, static_cast<size_t>(ceil_div(C, kScanChunk) + 1) * 4));
    MDX_HIP(hipMalloc(&partials, static_cast<size_t>(std::max(TB, 1)) * 3 * sizeof(double)));
    MDX_HIP(hipMalloc(&d_sp, nslots * sizeof(SysParams)));
    MDX_HIP(hipMalloc(&d_sc, nslots * sizeof(DevScalars)));
    MDX_HIP(hipMalloc(&d_blk2sys, std::max<size_t>(b2s.size(), 1) * 4));
    MDX_HIP(hipMalloc(&d_cblk2sys, std::max<size_t>(cb2s.size(), 1) * 4));
    MDX_HIP(hipMalloc(&d_lb2sys, std::max<size_t>(lb2s.size(), 1) * 4));
    MDX_HIP(hipMalloc(&d_actblk, std::max<size_t>(b2s.size(), 1) * 4));
    MDX_HIP(hipMalloc(&d_actcblk, std::max<size_t>(cb2s.size(), 1) * 4));
    MDX_HIP(hipMalloc(&d_want, nslots * 4));
    MDX_HIP(hipMalloc(&pid, n4));
    MDX_HIP(hipMalloc(&sid, n4));
    MDX_HIP(hipMemset(count, 0, (static_cast<size_t>(C) + 1) * 4));
    MDX_HIP(hipMemset(d_sc, 0, nslots * sizeof(DevScalars)));
    MDX_HIP(hipMemset(x, 0, n16));
    MDX_HIP(hipMemset(v, 0, n16));
    MDX_HIP(hipMemset(a, 0, n16));
    if (!b2s.empty()) MDX_HIP(hipMemcpy(d_blk2sys, b2s.data(), b2s.size() * 4, hipMemcpyHostToDevice));
    if (!cb2s.empty()) MDX_HIP(hipMemcpy(d_cblk2sys, cb2s.data(), cb2s.size() * 4, hipMemcpyHostToDevice));
    if (!lb2s.empty()) MDX_HIP(hipMemcpy(d_lb2sys, lb2s.data(), lb2s.size() * 4, hipMemcpyHostToDevice));
    MDX_HIP(hipMemcpy(d_sp, h_sp.data(), nslots * sizeof(SysParams), hipMemcpyHostToDevice));
    sp_dev = h_sp;
    ptrs.x = x;
    ptrs.v = v;
    ptrs.a = a;
    ptrs.cls = cls;
    ptrs.key = key;
    ptrs.count = count;
    ptrs.cs = cs;
    ptrs.sorted = sorted;
    ptrs.scell = scell;
    ptrs.img = img;
    ptrs.icls = icls;
    ptrs.partials = partials;
    ptrs.sorted_old = sorted_old;
    ptrs.cs_old = cs_old;
    ptrs.movers = movers;
    ptrs.mover_off = mover_off;
    ptrs.n_movers = n_movers;
    ptrs.arrivals = arrivals;
    ptrs.sp = d_sp;
    ptrs.sc = d_sc;
    ptrs.blk2sys = d_blk2sys;
    ptrs.actblk = d_actblk;
    ptrs.actcblk = d_actcblk;
    ptrs.cblk2sys = d_cblk2sys;
    ptrs.lb2sys = d_lb2sys;
    ptrs.pid = pid;
    ptrs.sid = sid;
    if (N > 0) {
        k_ident_ids<<<ceil_div(N, kBlock), kBlock, 0, stream>>>(pid, sid, 0u, static_cast<uint32_t>(N));
        MDX_HIP(hipGetLastError());
    }
    ptrs.TB = TB;
    ptrs.S = nslots;
    ptrs.C = C;
    ptrs.N = static_cast<int32_t>(N);
    return MDX_OK;
}

// SysParams go to the device only when they changed (a run of short step
// calls uploads nothing); pending pair partials are combined first, under
// the pair mode they were produced with.
int Batch::upload_sp(int r) {
    if (sp_dev.size() == h_sp.size() && std::memcmp(&sp_dev[r], &h_sp[r], sizeof(SysParams)) == 0) return MDX_OK;
    MDX_TRY(flush());
    MDX_HIP(hipMemcpyAsync(d_sp + r, &h_sp[r], sizeof(SysParams), hipMemcpyHostToDevice, stream));
    if (sp_dev.size() == h_sp.size()) sp_dev[r] = h_sp[r];
    return MDX_OK;
}

int Batch::upload_all_sp() {
    if (sp_dev.size() == h_sp.size() &&
        std::memcmp(sp_dev.data(), h_sp.data(), h_sp.size() * sizeof(SysParams)) == 0)
        return MDX_OK;
    MDX_TRY(flush());
    MDX_HIP(hipMemcpyAsync(d_sp, h_sp.data(), h_sp.size() * sizeof(SysParams), hipMemcpyHostToDevice, stream));
    sp_dev = h_sp;
    return MDX_OK;
}

int Batch::flush() {
    if (!fin_pending) return MDX_OK;
    fin_pending = false;
    k_finalize<<<static_cast<int>(slots.size()), kBlock, 0, stream>>>(ptrs, FIN_PAIR, d_want);
    MDX_HIP(hipGetLastError());
    return MDX_OK;
}

int Batch::read_scalars(int r, DevScalars &h) {
    MDX_TRY(flush());
    MDX_HIP(hipMemcpyAsync(&h, d_sc + r, sizeof(DevScalars), hipMemcpyDeviceToHost, stream));
    MDX_HIP(hipStreamSynchronize(stream));
    return MDX_OK;
}

int Batch::write_scalar_block(int r, const DevScalars &h) {
    MDX_TRY(flush());
    MDX_HIP(hipMemcpyAsync(d_sc + r, &h, sizeof(DevScalars), hipMemcpyHostToDevice, stream));
    MDX_HIP(hipStreamSynchronize(stream));
    return MDX_OK;
}

// Violation flags of the lean steps since the last check (one device-to-host
// read of every slot's scalars); the flags are cleared.
int Batch::check_violations() {
    viol_unchecked = false;
    const int S = static_cast<int>(slots.size());
    std::vector<DevScalars> cur(S);
    MDX_HIP(hipMemcpyAsync(cur.data(), d_sc, S * sizeof(DevScalars), hipMemcpyDeviceToHost, stream));
    MDX_HIP(hipStreamSynchronize(stream));
    bool bad = false;
    for (int r = 0; r < S; ++r) bad = bad || (slots[r].loaded && cur[r].violated != 0);
    if (!bad) return MDX_OK;
    k_clear_violations<<<ceil_div(S, kBlock), kBlock, 0, stream>>>(ptrs);
    MDX_HIP(hipGetLastError());
    return set_error(MDX_ERR_STATE,
                     "rebuild window: a particle moved beyond the reach of the lean-step fallback "
                     "(or the grid has fewer than 5 cells per side); use a shorter window");
}

int Batch::sync() {
    MDX_TRY(flush());
    MDX_HIP(hipStreamSynchronize(stream));
    if (viol_unchecked) return check_violations();
    return MDX_OK;
}

namespace {
// The neighbour-list parameters of a slot under the batch's tuning, set
// wherever a pass may run the slot in list mode (steps, FIRE, energy passes):
// a list built with unset ones (r_list 0) takes every candidate, overflows and
// turns list mode off for the slot.
void list_fields(SysParams &sp, const Slot &s, double skin, int list_cap) {
    sp.list_cap = list_cap;
    sp.skin_half2 = 0.25 * skin * skin;
    sp.r_list = s.rc_max + skin;
}

void set_idle(SysParams &sp, const Slot &s) {
    sp.kind = SK_NONE;
    sp.pmode = PM_NONE;
    sp.list_mode = 0;
    sp.periodic_y = s.cells_py;
    sp.wrap_keys = 1;
    sp.middle_wrap = 0;
}
}  // namespace

// Launch of a pair kernel; with events e0/e1, through hipExtLaunchKernel,
// whose start and stop events take the kernel's own start and end times (no
// dispatch overhead between them).
template <typename K>
static void launch_k(K kern, int grid, int block, hipStream_t st, hipEvent_t e0, hipEvent_t e1,
                     const BatchPtrs &p, int parity) {
    if (e0 && e1)
        hipExtLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, st, e0, e1, 0, p, static_cast<int32_t>(parity));
    else
        kern<<<grid, block, 0, st>>>(p, parity);
}

// FP32 pair kernel with lpp lanes per particle (1, 2 or 4).
template <int ACC, int OBS, bool UNI>
static void launch_pair_f32(int lpp, int grid, hipStream_t st, const BatchPtrs &p, int parity, int lean,
                            hipEvent_t e0, hipEvent_t e1) {
    if constexpr (OBS == SUMS_NONE) {   // lean steps take no sums
        if (lean) {
            switch (lpp) {
            case 2: launch_k(k_pair<int2, ACC, OBS, UNI, 2, true>, grid, kBlock * 2, st, e0, e1, p, parity); return;
            case 4: launch_k(k_pair<int2, ACC, OBS, UNI, 4, true>, grid, kBlock * 4, st, e0, e1, p, parity); return;
            default: launch_k(k_pair<int2, ACC, OBS, UNI, 1, true>, grid, kBlock, st, e0, e1, p, parity); return;
            }
        }
    }
    switch (lpp) {
    case 2: launch_k(k_pair<int2, ACC, OBS, UNI, 2, false>, grid, kBlock * 2, st, e0, e1, p, 0); break;
    case 4: launch_k(k_pair<int2, ACC, OBS, UNI, 4, false>, grid, kBlock * 4, st, e0, e1, p, 0); break;
    default: launch_k(k_pair<int2, ACC, OBS, UNI, 1, false>, grid, kBlock, st, e0, e1, p, 0); break;
    }
}

// Launch one step of every slot (their kind/pmode in d_sp).  A lean step
// (STEP_LEAN, step parity `parity`) is k_lean and the pair pass; every other
// step starts with k_box and runs the cell-list kernels.
int Batch::launch_step_kernels(int obs, int mode, int parity) {
    const int S = static_cast<int>(slots.size());
    const bool lean = mode == STEP_LEAN;
    if (lean && obs) return set_error(MDX_ERR_STATE, "a lean step takes no energy sums");
    if (!lean) parity = 0;
    // a window's head step keeps lists only if they last its window - 1 lean steps
    const int ahead = mode == STEP_HEAD ? std::max(window - 1, 0) : 0;
    if (!lean) k_box<<<S, kBlock, 0, stream>>>(ptrs, mode);
    if (TLB > 0 && lean) {   // windows run in the FP32 forms only
        if (prec == MDX_PREC_COMP32) k_lean<float2><<<TLB, kLeanThreads, 0, stream>>>(ptrs, parity);
        else k_lean<double2><<<TLB, kLeanThreads, 0, stream>>>(ptrs, parity);
    } else if (TA > 0 && !lean) {
        if (prec == MDX_PREC_COMP32)
            k_drift<float2><<<TA, kBlock, 0, stream>>>(ptrs, ahead);
        else
            k_drift<double2><<<TA, kBlock, 0, stream>>>(ptrs, ahead);
    }
    if (!lean) {   // cell list (and, in list mode, neighbour lists) of the slots that rebuild
        if (TCA + TA > 0) k_stay_arrive<<<TCA + TA, kBlock, 0, stream>>>(ptrs, TCA);
        const int nchunk = ceil_div(C, kScanChunk);
        if (nchunk > 0) {
            k_scan_a<<<nchunk, kBlock, 0, stream>>>(ptrs, C, bsum);
            k_scan_b<<<nchunk, kBlock, 0, stream>>>(ptrs, C, bsum);
        }
        if (TCA + TA > 0) k_fill<<<TCA + TA, kBlock, 0, stream>>>(ptrs, TCA);
        if (TCA > 0) k_cell_sort<<<TCA, kBlock, 0, stream>>>(ptrs);
    }
    if (TA > 0 && !lean) {
        if (prec == MDX_PREC_FP64)
            k_slot_image<double2><<<TA, kBlock, 0, stream>>>(ptrs);
        else
            k_slot_image<int2><<<TA, kBlock, 0, stream>>>(ptrs);
        if (nlist && prec != MDX_PREC_FP64) k_list_build<<<TA, kBlock, 0, stream>>>(ptrs);
    }
    if (TA > 0) {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (timing && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) {
            if (e0) (void)hipEventDestroy(e0);
            e0 = e1 = nullptr;
        }
        launch_pair(obs, parity, lean, e0, e1);
        if (e0 && e1) force_ev.emplace_back(e0, e1);
    }
    MDX_HIP(hipGetLastError());
    return MDX_OK;
}

// The pair kernel of the batch's precision, lanes and sigma form, with the
// sums obs (PairSums; the FP64 form takes its energy sums for the virial);
// lean: a rebuild window's lean-step variant with step parity `parity`;
// e0/e1: events for the kernel's start and end (launch_k), or null.
void Batch::launch_pair(int obs, int parity, bool lean, hipEvent_t e0, hipEvent_t e1) {
    const bool uni = uniform_sigma();
    const int ln = lean ? 1 : 0;
    switch (prec) {
    case MDX_PREC_COMP32:
        if (obs == SUMS_ENERGY) {
            if (uni) launch_pair_f32<ACC_COMP32, SUMS_ENERGY, true>(lpp, TA, stream, ptrs, parity, ln, e0, e1);
            else launch_pair_f32<ACC_COMP32, SUMS_ENERGY, false>(lpp, TA, stream, ptrs, parity, ln, e0, e1);
        } else if (obs == SUMS_VIRIAL) {
            if (uni) launch_pair_f32<ACC_COMP32, SUMS_VIRIAL, true>(lpp, TA, stream, ptrs, parity, ln, e0, e1);
            else launch_pair_f32<ACC_COMP32, SUMS_VIRIAL, false>(lpp, TA, stream, ptrs, parity, ln, e0, e1);
        } else {
            if (uni) launch_pair_f32<ACC_COMP32, SUMS_NONE, true>(lpp, TA, stream, ptrs, parity, ln, e0, e1);
            else launch_pair_f32<ACC_COMP32, SUMS_NONE, false>(lpp, TA, stream, ptrs, parity, ln, e0, e1);
        }
        break;
    case MDX_PREC_ACC64:
        if (obs == SUMS_ENERGY) {
            if (uni) launch_pair_f32<ACC_FP64, SUMS_ENERGY, true>(lpp, TA, stream, ptrs, parity, ln, e0, e1);
            else launch_pair_f32<ACC_FP64, SUMS_ENERGY, false>(lpp, TA, stream, ptrs, parity, ln, e0, e1);
        } else if (obs == SUMS_VIRIAL) {
            if (uni) launch_pair_f32<ACC_FP64, SUMS_VIRIAL, true>(lpp, TA, stream, ptrs, parity, ln, e0, e1);
            else launch_pair_f32<ACC_FP64, SUMS_VIRIAL, false>(lpp, TA, stream, ptrs, parity, ln, e0, e1);
        } else {
            if (uni) launch_pair_f32<ACC_FP64, SUMS_NONE, true>(lpp, TA, stream, ptrs, parity, ln, e0, e1);
            else launch_pair_f32<ACC_FP64, SUMS_NONE, false>(lpp, TA, stream, ptrs, parity, ln, e0, e1);
        }
        break;
    default:   // the FP64 form keeps one lane per particle (the oracle's order) and its FP64 energy sums
        if (obs != SUMS_NONE)
            launch_k(k_pair<double2, ACC_FP64, SUMS_ENERGY, false, 1, false>, TA, kBlock, stream, e0, e1, ptrs, 0);
        else
            launch_k(k_pair<double2, ACC_FP64, SUMS_NONE, false, 1, false>, TA, kBlock, stream, e0, e1, ptrs, 0);
        break;
    }
}

// Duration of the pair pass a step of the current tuning runs (the lean-step
// variant when rebuild windows apply), without launch overhead: reps launches
// of it captured back to back in one graph, timed with HIP events around the
// graph's launch on the batch's stream.  The launches run on the current
// state; the arrays and scalars they write (velocities, accelerations, the
// reduction partials, the slots' scalars) are saved before and restored
// after, so the state is unchanged.
int Batch::time_pair_pass(int reps, double *avg_ms) {
    if (reps < 1 || reps > 4096) return set_error(MDX_ERR_INVALID, "reps must be in [1, 4096]");
    if (TA <= 0) return set_error(MDX_ERR_STATE, "no particles");
    MDX_HIP(hipSetDevice(device));
    MDX_TRY(flush());
    MDX_HIP(hipStreamSynchronize(stream));
    const bool lean = windows_apply();
    const size_t nv = static_cast<size_t>(std::max<int64_t>(N, 1)) * sizeof(double2);
    const size_t na = static_cast<size_t>(std::max<int64_t>(N, 1)) * (prec == MDX_PREC_COMP32 ? sizeof(float2)
                                                                                               : sizeof(double2));
    const size_t np = static_cast<size_t>(TB) * 3 * sizeof(double);
    const size_t ns = slots.size() * sizeof(DevScalars);
    void *save = nullptr;
    MDX_HIP(hipMalloc(&save, nv + na + np + ns));
    char *sv = static_cast<char *>(save);
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = MDX_OK;
    auto hip = [&](hipError_t e) {
        if (e != hipSuccess && rc == MDX_OK) rc = set_error(MDX_ERR_HIP, hipGetErrorString(e));
        return e == hipSuccess;
    };
    if (hip(hipMemcpyAsync(sv, v, nv, hipMemcpyDeviceToDevice, stream)) &&
        hip(hipMemcpyAsync(sv + nv, a, na, hipMemcpyDeviceToDevice, stream)) &&
        hip(hipMemcpyAsync(sv + nv + na, partials, np, hipMemcpyDeviceToDevice, stream)) &&
        hip(hipMemcpyAsync(sv + nv + na + np, d_sc, ns, hipMemcpyDeviceToDevice, stream)) &&
        hip(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal))) {
        for (int i = 0; i < reps; ++i) launch_pair(SUMS_NONE, 0, lean, nullptr, nullptr);
        const bool ended = hip(hipStreamEndCapture(stream, &g));
        if (ended && hip(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0)) && hip(hipEventCreate(&e0)) &&
            hip(hipEventCreate(&e1)) && hip(hipEventRecord(e0, stream)) && hip(hipGraphLaunch(ge, stream)) &&
            hip(hipEventRecord(e1, stream)) && hip(hipEventSynchronize(e1))) {
            float ms = 0.0f;
            if (hip(hipEventElapsedTime(&ms, e0, e1))) *avg_ms = static_cast<double>(ms) / reps;
        }
    }
    // the saved state back whatever happened above
    (void)hipMemcpyAsync(v, sv, nv, hipMemcpyDeviceToDevice, stream);
    (void)hipMemcpyAsync(a, sv + nv, na, hipMemcpyDeviceToDevice, stream);
    (void)hipMemcpyAsync(partials, sv + nv + na, np, hipMemcpyDeviceToDevice, stream);
    (void)hipMemcpyAsync(d_sc, sv + nv + na + np, ns, hipMemcpyDeviceToDevice, stream);
    hip(hipStreamSynchronize(stream));
    if (ge) (void)hipGraphExecDestroy(ge);
    if (g) (void)hipGraphDestroy(g);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(save);
    return rc;
}

// count passes of the step sequence; replayed from a captured hipGraph unless
// kernel timing is on.
bool Batch::list_kind(int kind) const {
    return nlist && skin > 0.0 && prec != MDX_PREC_FP64 &&
           (kind == SK_NVE || kind == SK_NVT || kind == SK_NPH || kind == SK_FIRE);
}

bool Batch::uniform_sigma() const {
    for (const Slot &s : slots)
        if (s.loaded && !(s.p.sigma_aa == s.p.sigma_bb && s.p.sigma_aa == s.p.sigma_ab)) return false;
    return true;
}

// count steps of one mode, replayed from a captured hipGraph unless kernel
// timing is on.
// Captures `steps` steps of one mode and parity into ge (when it is empty).
int Batch::capture_steps(hipGraphExec_t &ge, int obs, int mode, int parity, int steps) {
    if (ge) return MDX_OK;
    hipGraph_t g = nullptr;
    MDX_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    int rc = MDX_OK;
    for (int k = 0; k < steps && rc == MDX_OK; ++k) rc = launch_step_kernels(obs, mode, parity);
    hipError_t e = hipStreamEndCapture(stream, &g);
    if (rc != MDX_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    MDX_HIP(e);
    e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    MDX_HIP(e);
    return MDX_OK;
}

int Batch::replay(int obs, int mode, int count_, int parity) {
    if (count_ <= 0) return MDX_OK;
    if (mode != STEP_LEAN) parity = 0;
    if (use_graph && !timing) {
        const int oi = obs;
        int k = 0;
        if (count_ >= kGraphSteps && mode != STEP_LEAN) {   // fewer launches: kGraphSteps steps per graph
            hipGraphExec_t &gm = graph_multi[oi][mode];
            MDX_TRY(capture_steps(gm, obs, mode, 0, kGraphSteps));
            for (; k + kGraphSteps <= count_; k += kGraphSteps) MDX_HIP(hipGraphLaunch(gm, stream));
        }
        if (k < count_) {
            hipGraphExec_t &ge = graph[oi][mode][parity];
            MDX_TRY(capture_steps(ge, obs, mode, parity, 1));
            for (; k < count_; ++k) MDX_HIP(hipGraphLaunch(ge, stream));
        }
    } else {
        for (int k = 0; k < count_; ++k) MDX_TRY(launch_step_kernels(obs, mode, parity));
    }
    return MDX_OK;
}

// Windows apply when the batch is tuned for them and every stepping slot is a
// list-mode slot (NVE / NVT in an FP32 form).
bool Batch::windows_apply() {
    if (window < 2 || !nlist || prec == MDX_PREC_FP64) return false;
    bool any = false;
    for (size_t r = 0; r < slots.size(); ++r) {
        Slot &sl = slots[r];
        if (!sl.loaded || h_sp[r].kind == SK_NONE) continue;
        // lean steps are NVE / NVT steps (an NPH slot in list mode rebuilds on demand)
        if (!h_sp[r].list_mode || (h_sp[r].kind != SK_NVE && h_sp[r].kind != SK_NVT)) return false;
        // the lean-step fallback needs 5 distinct cells per side; list-mode
        // kinds (NVE, NVT) keep the box, which is read back once after an NPH
        // or piston run may have changed it
        if (sl.box_moved) {
            DevScalars h;
            if (read_scalars(static_cast<int>(r), h) != MDX_OK) return false;
            sl.Lx_now = h.g.Lx;
            sl.Ly_now = h.g.Ly;
            sl.box_moved = false;
        }
        const Grid g = make_grid(sl.Lx_now, sl.Ly_now, h_sp[r].cell_side, sl.rc_max, 1);
        if (g.ncx < 5 || g.ncy < 5) return false;
        any = true;
    }
    return any;
}

// count passes of the step sequence; a run ends with the combination of the
// last pair pass's reduction (earlier ones are combined by the next k_box).
int Batch::run_passes(int count_, int obs) {
    if (count_ <= 0) return MDX_OK;
    if (graph_uni != uniform_sigma()) {   // captured graphs hold the other pair-kernel variant
        invalidate_graphs();
        graph_uni = uniform_sigma();
    }
    hipEvent_t s0 = nullptr, s1 = nullptr;
    if (timing && hipEventCreate(&s0) == hipSuccess && hipEventCreate(&s1) == hipSuccess)
        (void)hipEventRecord(s0, stream);
    if (!obs && windows_apply()) {
        MDX_TRY(run_windows(count_));
    } else {
        win_pos = 0;
        MDX_TRY(replay(obs, STEP_NORMAL, count_));
        k_finalize<<<static_cast<int>(slots.size()), kBlock, 0, stream>>>(ptrs, FIN_PAIR, d_want);
        MDX_HIP(hipGetLastError());
        fin_pending = false;
    }
    if (s1) {
        (void)hipEventRecord(s1, stream);
        step_ev.emplace_back(s0, s1);
    }
    return MDX_OK;
}

// Rebuild windows (list mode): each window is a head step (STEP_HEAD: the
// slot rebuilds its lists unless they last the window's lean steps) and
// window - 1 lean steps (no cell-list kernels; a slot with a particle beyond
// skin / 2 anyway, or without a usable list, takes the 5x5 pass over its last
// cell list, which is exact while every particle stays within (2 w - rc_max)
// / 2 of that build).  Windows run on across calls: a call first finishes
// the window the last call left open (win_pos), so the rebuild points and
// results do not depend on how the steps are split into calls.  No host
// synchronisation: the last pair pass's partials are left for the next
// kernel that combines them (fin_pending), and the slots' violation flags are
// read at the end of a long call (at least kWindowCheckSteps steps) or at the
// next synchronising call (Batch::sync, observables, collect).
int Batch::run_windows(int count_) {
    int left = count_;
    while (win_pos > 0 && left > 0) {   // the open window's remaining lean steps
        MDX_TRY(replay(SUMS_NONE, STEP_LEAN, 1, win_pos & 1));
        win_pos = (win_pos + 1) % window;
        --left;
    }
    while (left > 0) {
        const int w = std::min(window, left);
        if (w == window && use_graph && !timing) {
            // a whole window as one graph: one launch per window instead of
            // one per step, which is what small systems wait for
            if (!graph_window) {
                hipGraph_t g = nullptr;
                MDX_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
                int rc = launch_step_kernels(SUMS_NONE, STEP_HEAD, 0);
                for (int j = 1; j < w && rc == MDX_OK; ++j) rc = launch_step_kernels(SUMS_NONE, STEP_LEAN, j & 1);
                hipError_t e = hipStreamEndCapture(stream, &g);
                if (rc != MDX_OK) {
                    if (g) (void)hipGraphDestroy(g);
                    return rc;
                }
                MDX_HIP(e);
                e = hipGraphInstantiate(&graph_window, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                MDX_HIP(e);
            }
            MDX_HIP(hipGraphLaunch(graph_window, stream));
        } else {
            MDX_TRY(replay(SUMS_NONE, STEP_HEAD, 1));   // step parity 0
            for (int j = 1; j < w; ++j) MDX_TRY(replay(SUMS_NONE, STEP_LEAN, 1, j & 1));
        }
        left -= w;
        win_pos = w % window;
    }
    fin_pending = true;
    viol_unchecked = true;
    if (count_ >= kWindowCheckSteps) return check_violations();
    return MDX_OK;
}


int Batch::load(int r, const mdx_params &p, const mdx_particle64 *frame, int64_t n) {
    if (r < 0 || r >= static_cast<int>(slots.size())) return set_error(MDX_ERR_INVALID, "slot out of range");
    if (!frame || n <= 0) return set_error(MDX_ERR_INVALID, "empty frame");
    Slot &s = slots[r];
    if (n > s.n_cap) return set_error(MDX_ERR_CAPACITY, "frame larger than the slot capacity");
    MDX_TRY(validate_params(p));
    win_pos = 0;
    MDX_HIP(hipSetDevice(device));
    SysParams &sp = h_sp[r];
    double rc_max = 0.0;
    sp.pt = make_pair_table(p, rc_max);
    const double side = std::max(p.cell_scale, 1.0) * rc_max;
    const Grid g = make_grid(p.Lx, p.Ly, side, rc_max, 1);
    if (static_cast<int64_t>(g.ncx) * g.ncy > s.ccap)
        return set_error(MDX_ERR_CAPACITY, "box needs more cells than the slot holds");
    // the cells the grid may reach while this job is loaded: the largest box
    // the barostat clamps allow (height up to barostat_height_max_ratio of
    // the loaded one; width through a changed aspect ratio up to max / min of
    // the ratios), at the smallest cell side a tuning sets (rc_max)
    const double fy = std::max(1.0, p.barostat_height_max_ratio);
    const double fx = fy / std::min(1.0, std::max(p.barostat_height_min_ratio, 1.0e-3));
    const Grid gmax = make_grid(p.Lx * fx, p.Ly * fy, rc_max, rc_max, 1);
    const int64_t cb = std::min<int64_t>(s.ccap, static_cast<int64_t>(gmax.ncx) * gmax.ncy);
    sp.cbound = static_cast<int32_t>(std::max<int64_t>(cb, static_cast<int64_t>(g.ncx) * g.ncy));
    s.p = p;
    s.rc_max = rc_max;
    s.n = n;
    s.loaded = true;
    s.t = 0.0;
    s.Lx0 = p.Lx;
    s.Ly0 = p.Ly;
    s.Lx_now = p.Lx;
    s.Ly_now = p.Ly;
    s.box_moved = false;
    s.cells_valid = false;
    s.cells_py = 1;
    s.K_valid = s.W_valid = false;
    s.piston_active = s.nph_init = s.vp_init = false;
    s.kind = SK_NONE;
    s.flags = 0;
    sp.n = static_cast<int32_t>(n);
    sp.cell_side = cell_side_for(s);
    sp.rc_max = rc_max;
    sp.dt = p.dt;
    sp.T0 = p.T_target;
    sp.Q = p.Q;
    sp.g_dof = 2.0 * static_cast<double>(n) - 2.0;   // src/md_env.cu:1667
    sp.W_baro = std::max(p.barostat_mass, 1e-12);
    sp.k_piston = p.k_piston;
    set_idle(sp, s);
    // particles
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
    MDX_HIP(hipMemcpyAsync(x + sp.off, hx.data(), n * 16, hipMemcpyHostToDevice, stream));
    MDX_HIP(hipMemcpyAsync(v + sp.off, hv.data(), n * 16, hipMemcpyHostToDevice, stream));
    MDX_HIP(hipMemcpyAsync(cls + sp.off, hc.data(), n, hipMemcpyHostToDevice, stream));
    MDX_TRY(reset_ids(r));
    DevScalars h{};
    h.g = g;
    h.lin_cum = 1.0;
    h.gamma = p.Lx / p.Ly;
    h.Ly_floor = std::max(1e-12, p.barostat_height_min_ratio * p.Ly);
    h.Ly_ceil = std::max(h.Ly_floor, p.barostat_height_max_ratio * p.Ly);
    MDX_HIP(hipMemcpyAsync(d_sc + r, &h, sizeof(h), hipMemcpyHostToDevice, stream));
    MDX_TRY(upload_sp(r));
    MDX_HIP(hipStreamSynchronize(stream));
    MDX_TRY(update_active());
    return compute_forces({r}, 1, false);   // src/md_env.cu:66
}

int Batch::unload(int r) {
    if (r < 0 || r >= static_cast<int>(slots.size())) return set_error(MDX_ERR_INVALID, "slot out of range");
    Slot &s = slots[r];
    win_pos = 0;
    s.loaded = false;
    s.n = 0;
    h_sp[r].n = 0;
    set_idle(h_sp[r], s);
    MDX_TRY(upload_sp(r));
    return update_active();
}

int Batch::set_state(int r, const mdx_particle64 *frame, int64_t n) {
    if (r < 0 || r >= static_cast<int>(slots.size()) || !slots[r].loaded)
        return set_error(MDX_ERR_STATE, "slot not loaded");
    Slot &s = slots[r];
    if (n != s.n)
        return set_error(MDX_ERR_INVALID, "loaded particle count " + std::to_string(n) + " doesn't match n " +
                                              std::to_string(s.n));
    const int64_t off = h_sp[r].off;
    win_pos = 0;
    MDX_TRY(flush());
    std::vector<double2> hx(n), hv(n);
    std::vector<uint8_t> hc(n);
    for (int64_t i = 0; i < n; ++i) {
        hx[i] = make_double2(frame[i].pos[0], frame[i].pos[1]);
        hv[i] = make_double2(frame[i].vel[0], frame[i].vel[1]);
        const int t = frame[i].type;
        hc[i] = static_cast<uint8_t>(t == 0 ? 0 : (t == 1 ? 1 : 2));
        s.type[i] = t;
        s.theta[i] = frame[i].theta;
    }
    MDX_HIP(hipMemcpyAsync(x + off, hx.data(), n * 16, hipMemcpyHostToDevice, stream));
    MDX_HIP(hipMemcpyAsync(v + off, hv.data(), n * 16, hipMemcpyHostToDevice, stream));
    MDX_HIP(hipMemcpyAsync(cls + off, hc.data(), n, hipMemcpyHostToDevice, stream));
    MDX_TRY(reset_ids(r));
    {   // the cell list names the old storage: the next pass builds it anew
        DevScalars h;
        MDX_TRY(read_scalars(r, h));
        h.list_valid = 0;
        h.list_built = 0;
        MDX_TRY(write_scalar_block(r, h));
    }
    MDX_HIP(hipStreamSynchronize(stream));
    s.piston_active = false;
    s.K_valid = s.W_valid = false;
    return compute_forces({r}, 1, false);
}

int Batch::set_kind(int r, int kind, double p_ext, uint32_t flags) {
    if (r < 0 || r >= static_cast<int>(slots.size())) return set_error(MDX_ERR_INVALID, "slot out of range");
    if (kind != SK_NONE && !slots[r].loaded) return set_error(MDX_ERR_STATE, "slot not loaded");
    // another kind of step ends the open rebuild window (the same kind, from
    // the next call, continues it)
    if (slots[r].kind != kind || slots[r].p_ext != p_ext || slots[r].flags != flags) win_pos = 0;
    slots[r].kind = kind;
    slots[r].p_ext = p_ext;
    slots[r].flags = flags;
    if (kind == SK_NPH || kind == SK_PISTON) slots[r].box_moved = true;
    return MDX_OK;
}

int Batch::kinetic(const std::vector<int> &which) {
    if (which.empty()) return MDX_OK;
    MDX_TRY(flush());   // k_kinetic writes the partials (a lean step would read them)
    win_pos = 0;
    std::fill(h_want.begin(), h_want.end(), 0);
    for (int r : which) h_want[r] = 1;
    MDX_HIP(hipMemcpyAsync(d_want, h_want.data(), h_want.size() * 4, hipMemcpyHostToDevice, stream));
    if (TB > 0) {
        k_kinetic<<<TB, kBlock, 0, stream>>>(ptrs, d_want);
        k_finalize<<<static_cast<int>(slots.size()), kBlock, 0, stream>>>(ptrs, FIN_KINETIC, d_want);
    }
    MDX_HIP(hipGetLastError());
    for (int r : which) slots[r].K_valid = true;
    return MDX_OK;
}

int Batch::walls(const std::vector<int> &which) {
    if (which.empty()) return MDX_OK;
    MDX_TRY(flush());
    win_pos = 0;
    std::fill(h_want.begin(), h_want.end(), 0);
    for (int r : which) h_want[r] = 1;
    MDX_HIP(hipMemcpyAsync(d_want, h_want.data(), h_want.size() * 4, hipMemcpyHostToDevice, stream));
    if (TB > 0) {
        k_walls<<<TB, kBlock, 0, stream>>>(ptrs, d_want);
        k_finalize<<<static_cast<int>(slots.size()), kBlock, 0, stream>>>(ptrs, FIN_WALLS, d_want);
    }
    MDX_HIP(hipGetLastError());
    return MDX_OK;
}

// One pass at the current positions: slots in `which` rebuild their cell list
// with the given y boundary and run pair mode pmode; the others idle.
// A pass at the current positions for the slots in `which`.  Energy passes
// (PM_OBS) of NVE / NVT slots in list mode use the neighbour lists and rebuild
// them, as a step does, only when a particle has moved beyond skin / 2 since
// the last build: so when energies are sampled does not move the rebuild
// points of the trajectory (the cell list is canonical).  Other passes rebuild
// the cell list at the current positions.
static int one_pass(Batch &b, const std::vector<int> &which, int periodic_y, int pmode) {
    for (size_t r = 0; r < b.slots.size(); ++r) set_idle(b.h_sp[r], b.slots[r]);
    for (int r : which) {
        SysParams &sp = b.h_sp[r];
        const Slot &sl = b.slots[r];
        sp.kind = SK_KEYS;
        sp.periodic_y = periodic_y;
        sp.pmode = pmode;
        sp.wrap_keys = 1;
        sp.list_mode = (pmode == PM_OBS && b.list_kind(sl.kind)) ? 1 : 0;
        list_fields(sp, sl, b.skin, b.list_cap);
    }
    MDX_TRY(b.upload_all_sp());
    MDX_TRY(b.run_passes(1, pmode == PM_OBS ? SUMS_ENERGY : SUMS_NONE));
    b.win_pos = 0;   // the next step starts a window (a head step) from this pass's cell list
    for (int r : which) {
        b.slots[r].cells_valid = true;
        b.slots[r].cells_py = periodic_y;
    }
    return MDX_OK;
}

int Batch::rebuild_cells(const std::vector<int> &which, int periodic_y) {
    return one_pass(*this, which, periodic_y, PM_NONE);
}

int Batch::compute_forces(const std::vector<int> &which, int periodic_y, bool with_walls) {
    MDX_TRY(one_pass(*this, which, periodic_y, with_walls ? PM_WALLS : PM_PLAIN));
    for (int r : which) slots[r].W_valid = false;
    return MDX_OK;
}

int Batch::obs_pass(int r, int periodic_y) { return one_pass(*this, {r}, periodic_y, PM_OBS); }

// Preconditions of the reference's steps that need a pass of their own:
// K for Nose-Hoover, P_old for NPH, cal_forces_piston on the first piston step.
int Batch::pre_step() {
    std::vector<int> need_k, need_w, need_piston;
    for (size_t r = 0; r < slots.size(); ++r) {
        Slot &s = slots[r];
        if (!s.loaded || s.kind == SK_NONE) continue;
        switch (s.kind) {
        case SK_NVE:
            s.piston_active = false;
            break;
        case SK_NVT:
            s.piston_active = false;
            if (!s.K_valid) need_k.push_back(static_cast<int>(r));
            break;
        case SK_NPH:
            s.piston_active = false;
            if (!s.nph_init) {   // init_nph_state_if_needed, src/md_env.cu:1757-1778
                DevScalars h;
                MDX_TRY(read_scalars(static_cast<int>(r), h));
                h.gamma = h.g.Lx / h.g.Ly;
                h.p_A = std::max(s.p.barostat_mass, 1e-12) * s.p.barostat_area_rate_init;
                MDX_TRY(write_scalar_block(static_cast<int>(r), h));
                s.nph_init = true;
            }
            if (!s.K_valid) need_k.push_back(static_cast<int>(r));
            if (!s.W_valid || !s.cells_valid || s.cells_py != 1) need_w.push_back(static_cast<int>(r));
            break;
        case SK_PISTON:
            if (!s.vp_init) {   // init_nph_piston_state_if_needed, src/md_env.cu:1780-1793
                DevScalars h;
                MDX_TRY(read_scalars(static_cast<int>(r), h));
                h.v_p = 0.0;
                MDX_TRY(write_scalar_block(static_cast<int>(r), h));
                s.vp_init = true;
            }
            if (!s.piston_active) {
                s.piston_active = true;
                need_piston.push_back(static_cast<int>(r));
            }
            break;
        default:
            break;
        }
    }
    MDX_TRY(kinetic(need_k));
    if (!need_w.empty()) {
        MDX_TRY(one_pass(*this, need_w, 1, PM_OBS));
        for (int r : need_w) slots[r].W_valid = true;
    }
    if (!need_piston.empty()) MDX_TRY(compute_forces(need_piston, 0, true));
    return MDX_OK;
}

int Batch::step(int nsteps) {
    if (nsteps < 0) return set_error(MDX_ERR_INVALID, "nsteps < 0");
    if (nsteps == 0) return MDX_OK;
    MDX_HIP(hipSetDevice(device));
    MDX_TRY(pre_step());
    int obs = SUMS_NONE;   // NPH slots take the pair pass's virial (the barostat's pressure)
    bool any = false, nph = false;
    for (size_t r = 0; r < slots.size(); ++r) {
        Slot &s = slots[r];
        SysParams &sp = h_sp[r];
        if (!s.loaded || s.kind == SK_NONE) {
            set_idle(sp, s);
            continue;
        }
        any = true;
        sp.kind = s.kind;
        sp.p_ext = s.p_ext;
        sp.wrap_keys = 0;
        sp.middle_wrap = 0;
        sp.periodic_y = s.kind == SK_PISTON ? 0 : 1;
        switch (s.kind) {
        case SK_NVE: sp.pmode = PM_NVE; break;
        case SK_NVT: sp.pmode = PM_NVT; break;
        case SK_NPH: sp.pmode = PM_NPH; obs = SUMS_VIRIAL; nph = true; break;
        case SK_PISTON: sp.pmode = PM_PISTON; break;
        case SK_FIRE: sp.pmode = PM_FIRE; break;
        default: return set_error(MDX_ERR_INVALID, "unknown ensemble");
        }
        sp.list_mode = list_kind(s.kind) ? 1 : 0;
        list_fields(sp, s, skin, list_cap);
        if (s.kind == SK_NVT && (s.flags & MDX_STEP_MIDDLE_WRAP)) {
            DevScalars h;
            MDX_TRY(read_scalars(static_cast<int>(r), h));
            const double Lx = h.g.Lx;
            if (Lx > 0.0 && h.g.Ly > 0.0 && s.p.devide_p > 0.0) {   // src/md_env.cu:1457-1473
                const double smax = std::max(s.p.sigma_aa, std::max(s.p.sigma_bb, s.p.sigma_ab));
                double slab = s.p.devide_p * Lx + 2.0 * smax;
                if (slab > Lx) slab = Lx;
                sp.middle_wrap = 1;
                sp.slab_width = slab;
                sp.slab_xmin = 0.5 * Lx - 0.5 * slab;
                sp.slab_period = 2.0 * slab;
            }
        }
    }
    if (!any) return MDX_OK;
    MDX_TRY(upload_all_sp());
    // a long run goes in chunks of reorder_every steps (with rebuild windows,
    // whole windows, so the windows fall as in one run); before each, the
    // slots that have run that many steps since their storage was put in cell
    // order are reordered (no effect on results).  A slot loaded or set since
    // is due at once: it takes one step (or window) for its first cell list
    for (int left = nsteps; left > 0;) {
        MDX_TRY(reorder_due());
        int chunk = left;
        if (reorder_every > 0) {
            chunk = std::min(left, reorder_every);
            for (const Slot &s : slots)
                if (s.loaded && s.kind != SK_NONE && !s.cells_valid && s.since_reorder >= reorder_every) chunk = 1;
        }
        if (chunk < left && windows_apply()) {
            // chunks end at window boundaries, where a reorder may run
            chunk = win_pos > 0 ? std::min(left, window - win_pos)
                                : std::min(left, std::max(window, chunk - chunk % window));
        }
        MDX_TRY(run_passes(chunk, obs));
        left -= chunk;
        for (Slot &s : slots)
            if (s.loaded && s.kind != SK_NONE) {
                s.since_reorder += chunk;
                s.cells_valid = true;   // every stepping kind leaves a complete cell list
            }
    }
    if (timing) step_n += nsteps;
    for (size_t r = 0; r < slots.size(); ++r) {
        Slot &s = slots[r];
        if (!s.loaded || s.kind == SK_NONE) continue;
        if (s.kind == SK_FIRE) {
            s.cells_valid = true;
            s.cells_py = 1;
            s.K_valid = s.W_valid = false;
            continue;
        }
        s.t += s.p.dt * nsteps;
        s.cells_valid = true;
        s.cells_py = s.kind == SK_PISTON ? 0 : 1;
        s.K_valid = s.kind == SK_NVT || s.kind == SK_NPH;
        s.W_valid = s.kind == SK_NPH;
    }
    if (nph) {   // the reference aborts on an invalid proposed area: every slot's flag in one read
        MDX_TRY(flush());
        std::vector<DevScalars> h(slots.size());
        MDX_HIP(hipMemcpyAsync(h.data(), d_sc, h.size() * sizeof(DevScalars), hipMemcpyDeviceToHost, stream));
        MDX_HIP(hipStreamSynchronize(stream));
        for (size_t r = 0; r < slots.size(); ++r)
            if (slots[r].loaded && slots[r].kind == SK_NPH && h[r].box_error)
                return set_error(MDX_ERR_BOX, "step_single_NPH: invalid proposed area");
    }
    return MDX_OK;
}

int Batch::observables(int r, uint32_t mask, mdx_obs *out) {
    if (r < 0 || r >= static_cast<int>(slots.size()) || !slots[r].loaded)
        return set_error(MDX_ERR_STATE, "slot not loaded");
    MDX_HIP(hipSetDevice(device));
    if (viol_unchecked) MDX_TRY(check_violations());   // a windowed run's deferred check
    Slot &s = slots[r];
    if ((mask & (MDX_OBS_K | MDX_OBS_P)) && !s.K_valid) MDX_TRY(kinetic({r}));
    if (s.piston_active) {
        if (mask & (MDX_OBS_U | MDX_OBS_VIRIAL)) MDX_TRY(obs_pass(r, 0));
        if (mask & (MDX_OBS_U | MDX_OBS_P)) MDX_TRY(walls({r}));
    } else if (mask & (MDX_OBS_U | MDX_OBS_P | MDX_OBS_VIRIAL)) {
        MDX_TRY(obs_pass(r, 1));
    }
    DevScalars h;
    MDX_TRY(read_scalars(r, h));
    std::memset(out, 0, sizeof(*out));
    out->K = h.K;
    const double Lx = h.g.Lx, Ly = h.g.Ly;
    if (mask & MDX_OBS_VIRIAL) out->W = h.W;
    if (s.piston_active) {
        if (mask & MDX_OBS_U) out->U = h.U + h.wallE;
        if (mask & MDX_OBS_P) {   // src/md_env.cu:2255-2261
            h.last_P = (Lx > 0.0) ? h.F_p / Lx : 0.0;
            MDX_TRY(write_scalar_block(r, h));
            out->P = h.last_P;
        }
    } else {
        if (mask & MDX_OBS_U) out->U = h.U;
        if (mask & MDX_OBS_P) {   // src/md_env.cu:2205-2252
            const double area = Lx * Ly;
            h.last_P = (area > 0.0) ? (2.0 * h.K + h.W) / (2.0 * area) : 0.0;
            MDX_TRY(write_scalar_block(r, h));
            out->P = h.last_P;
        }
    }
    out->xi = h.xi;
    out->p_A = h.p_A;
    out->area_rate = h.p_A / std::max(s.p.barostat_mass, 1e-12);
    out->v_p = h.v_p;
    out->F_p = h.F_p;
    out->last_P = h.last_P;
    out->Lx = Lx;
    out->Ly = Ly;
    out->t = s.t;
    out->piston_active = s.piston_active ? 1 : 0;
    out->eta = h.eta;
    return MDX_OK;
}

int Batch::partial_U_lambda(int r, double eps, double *out) {
    if (r < 0 || r >= static_cast<int>(slots.size()) || !slots[r].loaded)
        return set_error(MDX_ERR_STATE, "slot not loaded");
    MDX_HIP(hipSetDevice(device));
    MDX_TRY(obs_pass(r, slots[r].piston_active ? 0 : 1));
    DevScalars h;
    MDX_TRY(read_scalars(r, h));
    *out = -eps * h.lam;
    return MDX_OK;
}

int Batch::collect(int r, mdx_particle64 *out, int64_t cap) {
    if (r < 0 || r >= static_cast<int>(slots.size()) || !slots[r].loaded)
        return set_error(MDX_ERR_STATE, "slot not loaded");
    const Slot &s = slots[r];
    if (cap < s.n) return set_error(MDX_ERR_CAPACITY, "output capacity below n");
    MDX_HIP(hipSetDevice(device));
    if (viol_unchecked) MDX_TRY(check_violations());   // a windowed run's deferred check
    const int64_t n = s.n, off = h_sp[r].off;
    std::vector<double2> hx(n), hv(n);
    std::vector<unsigned char> ha(n * 16);
    const bool f32 = prec == MDX_PREC_COMP32;
    std::vector<uint32_t> hp(n);
    MDX_HIP(hipMemcpyAsync(hx.data(), x + off, n * 16, hipMemcpyDeviceToHost, stream));
    MDX_HIP(hipMemcpyAsync(hv.data(), v + off, n * 16, hipMemcpyDeviceToHost, stream));
    if (f32)
        MDX_HIP(hipMemcpyAsync(ha.data(), static_cast<float2 *>(a) + off, n * 8, hipMemcpyDeviceToHost, stream));
    else
        MDX_HIP(hipMemcpyAsync(ha.data(), static_cast<double2 *>(a) + off, n * 16, hipMemcpyDeviceToHost, stream));
    MDX_HIP(hipMemcpyAsync(hp.data(), pid + off, n * 4, hipMemcpyDeviceToHost, stream));
    MDX_HIP(hipStreamSynchronize(stream));
    // storage index i holds original particle pid[i] - off
    for (int64_t si = 0; si < n; ++si) {
        const int64_t i = static_cast<int64_t>(hp[si]) - off;
        if (i < 0 || i >= n) return set_error(MDX_ERR_STATE, "particle index map out of range");
        mdx_particle64 &q = out[i];
        std::memset(&q, 0, sizeof(q));
        q.pos[0] = hx[si].x;
        q.pos[1] = hx[si].y;
        q.vel[0] = hv[si].x;
        q.vel[1] = hv[si].y;
        if (f32) {
            float fa[2];
            std::memcpy(fa, &ha[si * 8], 8);
            q.acc[0] = fa[0];
            q.acc[1] = fa[1];
        } else {
            std::memcpy(q.acc, &ha[si * 16], 16);
        }
        q.type = s.type[i];
        q.theta = s.theta[i];
    }
    return MDX_OK;
}

int Batch::cell_list(int r, int periodic_y, int32_t *ncx, int32_t *ncy, uint32_t *keys, uint32_t *cell_start,
                     uint32_t *sorted_ids, int64_t cap_n, int64_t cap_cells) {
    if (r < 0 || r >= static_cast<int>(slots.size()) || !slots[r].loaded)
        return set_error(MDX_ERR_STATE, "slot not loaded");
    MDX_HIP(hipSetDevice(device));
    MDX_TRY(rebuild_cells({r}, periodic_y));
    DevScalars h;
    MDX_TRY(read_scalars(r, h));
    const Slot &s = slots[r];
    const SysParams &sp = h_sp[r];
    const int64_t ncell = static_cast<int64_t>(h.g.ncx) * h.g.ncy;
    if (ncx) *ncx = h.g.ncx;
    if (ncy) *ncy = h.g.ncy;
    if ((keys || sorted_ids) && cap_n < s.n) return set_error(MDX_ERR_CAPACITY, "cap_n below n");
    if (cell_start && cap_cells < ncell + 1) return set_error(MDX_ERR_CAPACITY, "cap_cells below ncell+1");
    std::vector<uint32_t> st(ncell + 1);
    MDX_HIP(hipMemcpyAsync(st.data(), cs + sp.coff, (ncell + 1) * 4, hipMemcpyDeviceToHost, stream));
    MDX_HIP(hipStreamSynchronize(stream));
    const uint32_t base = st[0];
    // device arrays are in storage order: map through pid to original indices
    std::vector<uint32_t> hp(s.n);
    MDX_HIP(hipMemcpyAsync(hp.data(), pid + sp.off, s.n * 4, hipMemcpyDeviceToHost, stream));
    MDX_HIP(hipStreamSynchronize(stream));
    const uint32_t off = static_cast<uint32_t>(sp.off);
    for (int64_t si = 0; si < s.n; ++si)
        if (hp[si] < off || hp[si] - off >= static_cast<uint64_t>(s.n))
            return set_error(MDX_ERR_STATE, "particle index map out of range");
    if (keys) {
        std::vector<uint32_t> hk(s.n);
        MDX_HIP(hipMemcpyAsync(hk.data(), key + sp.off, s.n * 4, hipMemcpyDeviceToHost, stream));
        MDX_HIP(hipStreamSynchronize(stream));
        for (int64_t si = 0; si < s.n; ++si) keys[hp[si] - off] = hk[si] - static_cast<uint32_t>(sp.coff);
    }
    if (cell_start)
        for (int64_t c = 0; c <= ncell; ++c) cell_start[c] = st[c] - base;
    if (sorted_ids) {
        std::vector<uint32_t> hs(s.n);
        MDX_HIP(hipMemcpyAsync(hs.data(), sorted + base, s.n * 4, hipMemcpyDeviceToHost, stream));
        MDX_HIP(hipStreamSynchronize(stream));
        for (int64_t k = 0; k < s.n; ++k) {
            const uint32_t si = hs[k] - off;
            if (si >= static_cast<uint64_t>(s.n)) return set_error(MDX_ERR_STATE, "cell list entry out of range");
            sorted_ids[k] = hp[si] - off;
        }
    }
    return MDX_OK;
}

int Batch::reset_ids(int r) {
    const SysParams &sp = h_sp[r];
    slots[r].since_reorder = kReorderDue;   // frames come in any order: reorder at the first chance
    if (sp.n_cap <= 0) return MDX_OK;
    k_ident_ids<<<ceil_div(sp.n_cap, kBlock), kBlock, 0, stream>>>(pid, sid, static_cast<uint32_t>(sp.off),
                                                                   static_cast<uint32_t>(sp.off + sp.n_cap));
    MDX_HIP(hipGetLastError());
    return MDX_OK;
}

namespace {
template <typename T>
void reorder_array(const Batch &b, T *arr) {
    T *tmp = static_cast<T *>(b.reorder_tmp);
    k_gather_cell_order<T><<<b.TB, kBlock, 0, b.stream>>>(b.ptrs, b.d_want, arr, tmp);
    k_copy_slots<T><<<b.TB, kBlock, 0, b.stream>>>(b.ptrs, b.d_want, tmp, arr);
}
}  // namespace

// Storage of the slots in `which` put in their current cell order: every
// per-particle array gathered through the cell list (sorted), which then maps
// cell-ordered slot k to storage k.  Passes read particle data through the
// cell list, so later accesses are coalesced; the cell order itself (by
// original index) and all reductions (in cell order) are unchanged, so
// results do not depend on when this runs.  Between steps only: the cell
// list is complete (list_valid) and sorted_old equals sorted.
int Batch::reorder(const std::vector<int> &which) {
    if (which.empty() || TB <= 0) return MDX_OK;
    if (!reorder_tmp) MDX_HIP(hipMalloc(&reorder_tmp, static_cast<size_t>(std::max<int64_t>(N, 1)) * 16));
    std::fill(h_want.begin(), h_want.end(), 0);
    for (int r : which) h_want[r] = 1;
    MDX_HIP(hipMemcpyAsync(d_want, h_want.data(), h_want.size() * 4, hipMemcpyHostToDevice, stream));
    reorder_array(*this, x);
    reorder_array(*this, v);
    if (prec == MDX_PREC_COMP32) reorder_array(*this, static_cast<float2 *>(a));
    else reorder_array(*this, static_cast<double2 *>(a));
    reorder_array(*this, cls);
    reorder_array(*this, key);
    reorder_array(*this, pid);
    if (key_list) reorder_array(*this, key_list);
    if (x0) reorder_array(*this, x0);
    k_reorder_ids<<<TB, kBlock, 0, stream>>>(ptrs, d_want);
    MDX_HIP(hipGetLastError());
    for (int r : which) slots[r].since_reorder = 0;
    return MDX_OK;
}

// Particle counts per x slab of slot r (k_n_profile), copied to out[2 nb].
int Batch::n_profile(int r, int nb, int32_t *out) {
    if (r < 0 || r >= static_cast<int>(slots.size()) || !slots[r].loaded)
        return set_error(MDX_ERR_STATE, "slot not loaded");
    if (nb < 1 || nb > kProfileMaxBins) return set_error(MDX_ERR_INVALID, "n_bins must be in [1, 8192]");
    if (!out) return set_error(MDX_ERR_INVALID, "null output");
    MDX_HIP(hipSetDevice(device));
    int32_t *d = nullptr;
    MDX_HIP(hipMalloc(&d, static_cast<size_t>(2 * nb) * sizeof(int32_t)));
    hipError_t e = hipMemsetAsync(d, 0, static_cast<size_t>(2 * nb) * sizeof(int32_t), stream);
    const int64_t n = slots[r].n;
    if (e == hipSuccess && n > 0) {
        k_n_profile<<<static_cast<int>((n + kBlock - 1) / kBlock), kBlock, static_cast<size_t>(2 * nb) * sizeof(int32_t),
                      stream>>>(ptrs, r, nb, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, d, static_cast<size_t>(2 * nb) * sizeof(int32_t), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    (void)hipFree(d);
    MDX_HIP(e);
    return MDX_OK;
}

// Pressure-tensor profile of slot r in nb slabs along y (get_pressure_profile,
// src/md_env.cu:3821-3918): a fresh cell list at the current positions (y
// periodic unless the slot's piston walls are active; the next step starts a
// new rebuild window), the fixed-point sums of k_pressure_profile, then each
// slab's sum divided by its area Lx * Ly / nb.  out: 3 nb values, Pxx for the
// nb slabs, then Pyy, then Pxy.
int Batch::pressure_profile(int r, int nb, double *out) {
    if (r < 0 || r >= static_cast<int>(slots.size()) || !slots[r].loaded)
        return set_error(MDX_ERR_STATE, "slot not loaded");
    if (nb < 1 || nb > kProfileMaxBins) return set_error(MDX_ERR_INVALID, "n_bins must be in [1, 8192]");
    if (!out) return set_error(MDX_ERR_INVALID, "null output");
    MDX_HIP(hipSetDevice(device));
    if (viol_unchecked) MDX_TRY(check_violations());
    MDX_TRY(rebuild_cells({r}, slots[r].piston_active ? 0 : 1));
    DevScalars h;
    MDX_TRY(read_scalars(r, h));
    const size_t bytes = static_cast<size_t>(3 * nb) * sizeof(unsigned long long);
    unsigned long long *d = nullptr;
    MDX_HIP(hipMalloc(&d, bytes));
    hipError_t e = hipMemsetAsync(d, 0, bytes, stream);
    const int64_t n = slots[r].n;
    constexpr int kLdsBins = 2048;   // 48 KB of LDS histogram
    if (e == hipSuccess && n > 0) {
        const int grid = static_cast<int>((n + kBlock - 1) / kBlock);
        if (nb <= kLdsBins)
            k_pressure_profile<true><<<grid, kBlock, bytes, stream>>>(ptrs, r, nb, d);
        else
            k_pressure_profile<false><<<grid, kBlock, 0, stream>>>(ptrs, r, nb, d);
        e = hipGetLastError();
    }
    std::vector<long long> hs(static_cast<size_t>(3 * nb));
    if (e == hipSuccess) e = hipMemcpyAsync(hs.data(), d, bytes, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    (void)hipFree(d);
    MDX_HIP(e);
    const double area = h.g.Lx * h.g.Ly / static_cast<double>(nb);
    for (int k = 0; k < 3 * nb; ++k) {
        const double v = static_cast<double>(hs[k]) / 4294967296.0;
        out[k] = area > 0.0 ? v / area : v;
    }
    return MDX_OK;
}

// middle_reflect_LG on slot r's current positions (no force pass, as in the
// reference); the next step rebuilds the cell list from the moved particles.
int Batch::middle_reflect(int r) {
    if (r < 0 || r >= static_cast<int>(slots.size()) || !slots[r].loaded)
        return set_error(MDX_ERR_STATE, "slot not loaded");
    MDX_HIP(hipSetDevice(device));
    Slot &s = slots[r];
    DevScalars h;
    MDX_TRY(read_scalars(r, h));
    const double Lx = h.g.Lx, Ly = h.g.Ly, dp = s.p.devide_p;
    if (Lx <= 0.0 || Ly <= 0.0 || dp <= 0.0) return MDX_OK;   // src/md_env.cu:1462-1464
    const double smax = std::max(s.p.sigma_aa, std::max(s.p.sigma_bb, s.p.sigma_ab));
    double slab = dp * Lx + 2.0 * smax;
    if (slab > Lx) slab = Lx;
    if (s.n > 0) {
        k_middle_reflect<<<ceil_div(s.n, kBlock), kBlock, 0, stream>>>(ptrs, r, 0.5 * Lx - 0.5 * slab, slab);
        MDX_HIP(hipGetLastError());
    }
    win_pos = 0;
    s.W_valid = false;
    MDX_HIP(hipStreamSynchronize(stream));
    return MDX_OK;
}

// ---- allocation-free N profile ------------------------------------------
// The counters of mdx_N_profile_noalloc: one fixed device array per GPU (each
// device that runs the library holds its own copy until the process ends),
// shared by every call on that device.  Host code never passes the symbol
// itself to a kernel or a copy: the device address comes from
// hipGetSymbolAddress, once per device, and is kept in scratch_ptr.
static __device__ __align__(256) int32_t scratch[2 * kProfileMaxBins];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
int32_t *g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's scratch array (the device must be current).
int scratch_ptr(int dev, int32_t **out) {
    if (dev < 0 || dev >= kMaxDevices) return set_error(MDX_ERR_INVALID, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(MDX_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<int32_t *>(p);
    }
    *out = g_scratch_ptr[dev];
    return MDX_OK;
}

__global__ void __launch_bounds__(kBlock) k_zero_i32(int32_t *p, int32_t n) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = 0;
}
}  // namespace

// n_profile with the counters in the device's scratch array instead of a
// buffer allocated per call: zeroed by a kernel (no memset node), filled by
// k_n_profile, copied to out.  Calls that overlap on one device race on the
// array.
int Batch::n_profile_noalloc(int r, int nb, int32_t *out) {
    if (r < 0 || r >= static_cast<int>(slots.size()) || !slots[r].loaded)
        return set_error(MDX_ERR_STATE, "slot not loaded");
    if (nb < 1 || nb > kProfileMaxBins)
        return set_error(MDX_ERR_CAPACITY, "n_bins must be in [1, " + std::to_string(kProfileMaxBins) +
                                               "] (the fixed scratch array holds 2 x " +
                                               std::to_string(kProfileMaxBins) + " counters)");
    if (!out) return set_error(MDX_ERR_INVALID, "null output");
    MDX_HIP(hipSetDevice(device));
    int32_t *d = nullptr;
    MDX_TRY(scratch_ptr(device, &d));
    const int n2 = 2 * nb;
    k_zero_i32<<<ceil_div(n2, kBlock), kBlock, 0, stream>>>(d, n2);
    MDX_HIP(hipGetLastError());
    const int64_t n = slots[r].n;
    if (n > 0) {
        k_n_profile<<<static_cast<int>((n + kBlock - 1) / kBlock), kBlock, static_cast<size_t>(n2) * sizeof(int32_t),
                      stream>>>(ptrs, r, nb, d);
        MDX_HIP(hipGetLastError());
    }
    MDX_HIP(hipMemcpyAsync(out, d, static_cast<size_t>(n2) * sizeof(int32_t), hipMemcpyDeviceToHost, stream));
    MDX_HIP(hipStreamSynchronize(stream));
    return MDX_OK;
}

int Batch::reorder_due() {
    if (win_pos != 0) return MDX_OK;   // not inside an open rebuild window
    std::vector<int> due;
    for (size_t r = 0; r < slots.size(); ++r) {
        const Slot &s = slots[r];
        if (reorder_every > 0 && s.loaded && s.cells_valid && s.since_reorder >= reorder_every)
            due.push_back(static_cast<int>(r));
    }
    return reorder(due);
}

int Batch::set_precision(int precision) {
    if (precision < MDX_PREC_COMP32 || precision > MDX_PREC_FP64)
        return set_error(MDX_ERR_INVALID, "unknown precision");
    if (precision == prec) return MDX_OK;
    MDX_HIP(hipSetDevice(device));
    MDX_HIP(hipStreamSynchronize(stream));
    prec = precision;
    invalidate_graphs();
    // accelerations and the cell image change representation: recompute
    std::vector<int> per, pis;
    for (size_t r = 0; r < slots.size(); ++r) {
        if (!slots[r].loaded) continue;
        (slots[r].piston_active ? pis : per).push_back(static_cast<int>(r));
    }
    if (!per.empty()) MDX_TRY(compute_forces(per, 1, false));
    if (!pis.empty()) MDX_TRY(compute_forces(pis, 0, true));
    MDX_HIP(hipStreamSynchronize(stream));
    return MDX_OK;
}

mdx_tuning Batch::tuning() const {
    mdx_tuning t{};
    t.cell_scale = cell_scale;
    t.skin = skin;
    t.lanes_per_particle = lpp;
    t.window = window;
    return t;
}

double Batch::cell_side_for(const Slot &s) const {
    const double side = std::max(s.p.cell_scale, 1.0) * s.rc_max;
    return skin > 0.0 ? std::max(side, s.rc_max + skin) : side;
}

int Batch::ensure_list_arrays() {
    if (nlist) return MDX_OK;
    // list entries carry the partner's slot in 27 bits (mdx_kernels.hpp, list_entry)
    if (N >= (int64_t{1} << 27)) return set_error(MDX_ERR_CAPACITY, "neighbour lists need fewer than 2^27 particles");
    if (list_cap % 4 != 0) return set_error(MDX_ERR_INVALID, "list capacity must be a multiple of 4");
    const size_t n = static_cast<size_t>(std::max<int64_t>(N, 1));
    MDX_HIP(hipMalloc(&key_list, n * 4));
    MDX_HIP(hipMalloc(&x0, n * sizeof(double2)));
    MDX_HIP(hipMalloc(&img0, n * sizeof(int2)));
    MDX_HIP(hipMalloc(&ncount, n * 4));
    MDX_HIP(hipMalloc(&nlist, n * static_cast<size_t>(list_cap) * 4));
    MDX_HIP(hipMemset(ncount, 0, n * 4));
    ptrs.key_list = key_list;
    ptrs.x0 = x0;
    ptrs.img0 = img0;
    ptrs.ncount = ncount;
    ptrs.nlist = nlist;
    invalidate_graphs();   // the step sequence gains the list build
    return MDX_OK;
}

int Batch::set_tuning(const mdx_tuning &t) {
    if (!(t.cell_scale >= 1.0) || !std::isfinite(t.cell_scale))
        return set_error(MDX_ERR_INVALID, "cell_scale must be finite and >= 1");
    if (t.lanes_per_particle != 1 && t.lanes_per_particle != 2 && t.lanes_per_particle != 4)
        return set_error(MDX_ERR_INVALID, "lanes_per_particle must be 1, 2 or 4");
    if (!(t.skin >= 0.0) || !(t.skin <= 4.0)) return set_error(MDX_ERR_INVALID, "skin must be in [0, 4]");
    if (t.window < 0 || t.window > 4096) return set_error(MDX_ERR_INVALID, "window must be in [0, 4096]");
    MDX_HIP(hipSetDevice(device));
    MDX_HIP(hipStreamSynchronize(stream));
    const int new_lpp = prec == MDX_PREC_FP64 ? 1 : t.lanes_per_particle;
    if (new_lpp != lpp) {
        lpp = new_lpp;
        invalidate_graphs();
    }
    const double new_skin = prec == MDX_PREC_FP64 ? 0.0 : t.skin;
    if (new_skin > 0.0) MDX_TRY(ensure_list_arrays());
    // the cell sides must fit the slots before anything changes
    for (size_t r = 0; r < slots.size(); ++r) {
        const Slot &s = slots[r];
        if (!s.loaded) continue;
        double side = std::max(t.cell_scale, 1.0) * s.rc_max;
        if (new_skin > 0.0) side = std::max(side, s.rc_max + new_skin);
        DevScalars h;
        MDX_TRY(read_scalars(static_cast<int>(r), h));
        const Grid g = make_grid(h.g.Lx, h.g.Ly, side, s.rc_max, 1);
        if (static_cast<int64_t>(g.ncx) * g.ncy > s.ccap)
            return set_error(MDX_ERR_CAPACITY, "cell_scale gives more cells than the slot holds");
    }
    cell_scale = t.cell_scale;
    skin = new_skin;
    const int new_window = new_skin > 0.0 ? t.window : 0;
    if (new_window != window) invalidate_graphs();   // the head step's look-ahead is captured
    window = new_window;
    std::vector<int> per, pis;
    for (size_t r = 0; r < slots.size(); ++r) {
        Slot &s = slots[r];
        if (!s.loaded) continue;
        s.p.cell_scale = t.cell_scale;
        h_sp[r].cell_side = cell_side_for(s);
        MDX_TRY(upload_sp(static_cast<int>(r)));
        DevScalars h;   // a list that overflowed gets another chance under the new tuning
        MDX_TRY(read_scalars(static_cast<int>(r), h));
        h.list_overflow = 0;
        // list_built stays: it names the keys the cell list was built from
        // (the movers' reference), and the force pass below rebuilds the cell
        // list, which clears it, so the next list-mode step builds new lists
        MDX_TRY(write_scalar_block(static_cast<int>(r), h));
        (s.piston_active ? pis : per).push_back(static_cast<int>(r));
    }
    // accelerations in the new summation order, as a fresh load would have them
    if (!per.empty()) MDX_TRY(compute_forces(per, 1, false));
    if (!pis.empty()) MDX_TRY(compute_forces(pis, 0, true));
    MDX_HIP(hipStreamSynchronize(stream));
    return MDX_OK;
}

int Batch::autotune(int steps, mdx_tune_report *rep, int kind_override) {
    if (steps <= 0) return set_error(MDX_ERR_INVALID, "autotune: steps must be positive");
    MDX_HIP(hipSetDevice(device));
    std::memset(rep, 0, sizeof(*rep));
    // a copy of the batch at the current state: same capacities and precision,
    // every loaded slot with its current frame, box and kind
    Batch c;
    std::vector<int64_t> caps;
    std::vector<int32_t> ccaps;
    for (const Slot &s : slots) {
        caps.push_back(s.n_cap);
        ccaps.push_back(s.ccap);
    }
    MDX_TRY(c.init(static_cast<int>(slots.size()), caps.data(), ccaps.data(), prec, device));
    bool any = false;
    for (size_t r = 0; r < slots.size(); ++r) {
        const Slot &s = slots[r];
        if (!s.loaded) continue;
        std::vector<mdx_particle64> f(s.n);
        MDX_TRY(collect(static_cast<int>(r), f.data(), s.n));
        DevScalars h;
        MDX_TRY(read_scalars(static_cast<int>(r), h));
        mdx_params p = s.p;
        p.Lx = h.g.Lx;
        p.Ly = h.g.Ly;
        MDX_TRY(c.load(static_cast<int>(r), p, f.data(), s.n));
        int k = kind_override >= 0 ? kind_override : s.kind;
        if (k == SK_NONE || k == SK_FIRE || k == SK_KEYS) k = SK_NVE;
        MDX_TRY(c.set_kind(static_cast<int>(r), k, s.p_ext, s.flags & ~MDX_STEP_MIDDLE_WRAP));
        any = true;
    }
    if (!any) return set_error(MDX_ERR_STATE, "autotune: no loaded slot");
    // candidates: cell list every step (cell size x lanes), or neighbour lists
    // (skin x lanes x rebuild window)
    std::vector<mdx_tuning> cands;
    for (double cs : {1.0, 1.1})
        for (int l : {1, 2, 4}) cands.push_back(mdx_tuning{cs, 0.0, l, 0});
    if (prec != MDX_PREC_FP64) {
        for (double sk : {0.3, 0.5, 0.65, 0.8})
            for (int w : {0, 8, 12, 16, 20}) cands.push_back(mdx_tuning{1.0, sk, 1, w});
        for (double sk : {0.5, 0.8}) cands.push_back(mdx_tuning{1.0, sk, 2, 0});
        // longer windows: a whole window is one graph launch
        cands.push_back(mdx_tuning{1.0, 0.8, 1, 24});
        cands.push_back(mdx_tuning{1.0, 1.0, 1, 20});
        cands.push_back(mdx_tuning{1.0, 1.0, 1, 24});
        cands.push_back(mdx_tuning{1.0, 1.0, 1, 32});
    }
    double best = 1e300;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    MDX_HIP(hipEventCreate(&e0));
    MDX_HIP(hipEventCreate(&e1));
    int rc = MDX_OK;
    for (const mdx_tuning &t : cands) {
        if (prec == MDX_PREC_FP64 && t.lanes_per_particle != 1) continue;
        if (rep->n_tried >= MDX_TUNE_MAX) break;
        rc = c.set_tuning(t);
        if (rc == MDX_ERR_CAPACITY) continue;   // too many cells for a slot: not a candidate
        if (rc != MDX_OK) break;
        // warm-up: every graph the timed steps launch is captured here (the
        // per-step graphs, a whole window's, eight steps'), and caches; with
        // windows, whole windows, so the timed steps start at a head step
        if (t.window >= 2) {
            if ((rc = c.step(2 * t.window)) != MDX_OK) break;
        } else {
            if ((rc = c.step(3)) != MDX_OK) break;
            if ((rc = c.step(kGraphSteps)) != MDX_OK) break;
        }
        // whole windows only, as in a long run
        const int ts = t.window >= 2 ? ((steps + t.window - 1) / t.window) * t.window : steps;
        (void)hipEventRecord(e0, c.stream);
        if ((rc = c.step(ts)) != MDX_OK) break;
        (void)hipEventRecord(e1, c.stream);
        if (hipEventSynchronize(e1) != hipSuccess) {
            rc = set_error(MDX_ERR_HIP, "autotune: event wait failed");
            break;
        }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        const double per = static_cast<double>(ms) / ts;
        rep->tried[rep->n_tried] = t;
        rep->ms_per_step[rep->n_tried] = per;
        ++rep->n_tried;
        if (per < best) {
            best = per;
            rep->best = t;
            rep->best_ms_per_step = per;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != MDX_OK && rc != MDX_ERR_CAPACITY) return rc;
    if (rep->n_tried == 0) return set_error(MDX_ERR_STATE, "autotune: no candidate fits the slots");
    return set_tuning(rep->best);
}

// Slot r starts FIRE (EnergyMinimizer::minimize_frame, src/md_energy_minimizer.cu:
// 122-160): zero velocities, dt = dt_init, alpha = alpha0, and the device
// schedule counters; its kind becomes SK_FIRE.
int Batch::fire_begin(int r, const mdx_fire_params &fp, double ftol, int max_steps) {
    if (r < 0 || r >= static_cast<int>(slots.size()) || !slots[r].loaded)
        return set_error(MDX_ERR_STATE, "slot not loaded");
    if (max_steps <= 0) return set_error(MDX_ERR_INVALID, "EnergyMinimizer: max_steps must be positive.");
    if (prec == MDX_PREC_COMP32)
        return set_error(MDX_ERR_INVALID, "FIRE slots need FP64 accelerations: use an ACC64 or FP64 batch");
    MDX_HIP(hipSetDevice(device));
    Slot &s = slots[r];
    s.piston_active = false;
    std::fill(h_want.begin(), h_want.end(), 0);
    h_want[r] = 1;
    MDX_HIP(hipMemcpyAsync(d_want, h_want.data(), h_want.size() * 4, hipMemcpyHostToDevice, stream));
    if (TB > 0) k_fire_init<<<TB, kBlock, 0, stream>>>(ptrs, d_want);   // :343-380
    MDX_HIP(hipGetLastError());
    DevScalars h;
    MDX_TRY(read_scalars(r, h));
    h.f_dt = fp.dt_init;
    h.f_alpha = fp.alpha0;
    h.f_maxF = h.f_power = 0.0;
    h.f_streak = h.f_done = h.f_zero_v = h.f_step = h.f_conv_step = h.f_drifted = 0;
    h.f_do_drift = h.f_do_pair = 0;
    MDX_TRY(write_scalar_block(r, h));
    SysParams &sp = h_sp[r];
    sp.ftol = ftol;
    sp.f_inc = fp.f_inc;
    sp.f_dec = fp.f_dec;
    sp.dt_max = fp.dt_max;
    sp.alpha0 = fp.alpha0;
    sp.N_min = fp.N_min;
    sp.max_step = fp.max_step;
    sp.f_max_steps = max_steps;
    s.kind = SK_FIRE;
    s.K_valid = s.W_valid = false;
    return upload_sp(r);
}

int Batch::fire_poll(int r, int *done, mdx_fire_result *res) {
    if (r < 0 || r >= static_cast<int>(slots.size()) || !slots[r].loaded)
        return set_error(MDX_ERR_STATE, "slot not loaded");
    MDX_HIP(hipSetDevice(device));
    DevScalars h;
    MDX_TRY(read_scalars(r, h));
    const int max_steps = h_sp[r].f_max_steps;
    const bool converged = h.f_done != 0;
    // unconverged: max_steps force passes and the drift after the last one
    const bool finished = converged || (h.f_step >= max_steps && h.f_drifted >= h.f_step);
    *done = finished ? 1 : 0;
    std::memset(res, 0, sizeof(*res));
    res->steps = converged ? h.f_conv_step : (finished ? max_steps : h.f_step);
    res->converged = converged ? 1 : 0;
    res->max_force = h.f_maxF;
    if (finished) {
        MDX_TRY(obs_pass(r, 1));   // :270, while the slot is a FIRE slot (its lists, as in Batch::fire)
        MDX_TRY(read_scalars(r, h));
        res->potential_energy = h.U;
        slots[r].kind = SK_NONE;
        set_idle(h_sp[r], slots[r]);
        MDX_TRY(upload_sp(r));
    }
    return MDX_OK;
}

int Batch::observables_many(const std::vector<int> &which, uint32_t mask, mdx_obs *out) {
    if (which.empty()) return MDX_OK;
    for (int r : which)
        if (r < 0 || r >= static_cast<int>(slots.size()) || !slots[r].loaded || slots[r].piston_active)
            return set_error(MDX_ERR_STATE, "observables_many: slot not loaded or in piston mode");
    MDX_HIP(hipSetDevice(device));
    if (mask & (MDX_OBS_K | MDX_OBS_P)) {
        std::vector<int> need;
        for (int r : which)
            if (!slots[r].K_valid) need.push_back(r);
        MDX_TRY(kinetic(need));
    }
    if (mask & (MDX_OBS_U | MDX_OBS_P | MDX_OBS_VIRIAL)) MDX_TRY(one_pass(*this, which, 1, PM_OBS));
    MDX_TRY(flush());
    std::vector<DevScalars> h(slots.size());
    MDX_HIP(hipMemcpyAsync(h.data(), d_sc, h.size() * sizeof(DevScalars), hipMemcpyDeviceToHost, stream));
    MDX_HIP(hipStreamSynchronize(stream));
    for (size_t q = 0; q < which.size(); ++q) {
        const int r = which[q];
        DevScalars &d = h[r];
        mdx_obs &o = out[q];
        std::memset(&o, 0, sizeof(o));
        const double area = d.g.Lx * d.g.Ly;
        o.K = d.K;
        if (mask & MDX_OBS_U) o.U = d.U;
        if (mask & MDX_OBS_VIRIAL) o.W = d.W;
        if (mask & MDX_OBS_P) {   // src/md_env.cu:2205-2252
            d.last_P = (area > 0.0) ? (2.0 * d.K + d.W) / (2.0 * area) : 0.0;
            o.P = d.last_P;
            MDX_HIP(hipMemcpyAsync(&d_sc[r].last_P, &d.last_P, sizeof(double), hipMemcpyHostToDevice, stream));
        }
        o.xi = d.xi;
        o.p_A = d.p_A;
        o.area_rate = d.p_A / std::max(slots[r].p.barostat_mass, 1e-12);
        o.v_p = d.v_p;
        o.F_p = d.F_p;
        o.last_P = d.last_P;
        o.Lx = d.g.Lx;
        o.Ly = d.g.Ly;
        o.t = slots[r].t;
        o.eta = d.eta;
    }
    MDX_HIP(hipStreamSynchronize(stream));
    return MDX_OK;
}

// FIRE on slot r (EnergyMinimizer::minimize_frame, src/md_energy_minimizer.cu:
// 122-272).  Standard step j runs the drift of iteration j-1 and the force
// pass of iteration j, so max_steps iterations take max_steps + 1 steps.
int Batch::fire(int r, const mdx_fire_params &fp, double ftol, int max_steps, mdx_fire_result *res,
                int32_t *es, double *et, int cap) {
    if (r < 0 || r >= static_cast<int>(slots.size()) || !slots[r].loaded)
        return set_error(MDX_ERR_STATE, "slot not loaded");
    if (max_steps <= 0) return set_error(MDX_ERR_INVALID, "EnergyMinimizer: max_steps must be positive.");
    MDX_HIP(hipSetDevice(device));
    Slot &s = slots[r];
    // Overlapping starts produce accelerations beyond FP32 range: FIRE keeps
    // FP64 accelerations (FP32 pairs with FP64 accumulation, or FP64 pairs).
    if (prec == MDX_PREC_COMP32) MDX_TRY(set_precision(MDX_PREC_ACC64));
    MDX_TRY(fire_begin(r, fp, ftol, max_steps));
    // this loop steps the slot itself (the others idle); its kind stays
    // SK_FIRE, so that its energy passes use the neighbour lists as its steps
    // do and leave their rebuild points where a run without them has them
    DevScalars h;
    SysParams &sp = h_sp[r];
    int ns = 0, last_pushed = -1;
    auto push = [&](int step, double U) {
        if (ns < cap) {
            es[ns] = step;
            et[ns] = U;
        }
        ++ns;
        last_pushed = step;
    };
    auto potential = [&](double &U) -> int {
        MDX_TRY(obs_pass(r, 1));
        DevScalars q;
        MDX_TRY(read_scalars(r, q));
        U = q.U;
        return MDX_OK;
    };
    double U = 0.0;
    MDX_TRY(potential(U));   // :145-147
    push(0, U);
    const int interval = fp.energy_interval > 0 ? fp.energy_interval : 1000;
    const int total = max_steps + 1;
    int j = 0;
    bool converged = false;
    while (j < total) {
        int end = std::min(total, ((j - 1 + interval) / interval) * interval + 1);
        if (end <= j) end = std::min(total, j + interval);
        end = std::min(end, j + 256);
        for (size_t q = 0; q < slots.size(); ++q) set_idle(h_sp[q], slots[q]);
        sp.kind = SK_FIRE;
        sp.pmode = PM_FIRE;
        sp.periodic_y = 1;
        sp.wrap_keys = 0;
        // neighbour lists (FP32 pair forms, skin > 0): rebuilt when an
        // iteration has moved some particle beyond skin / 2 since the build
        sp.list_mode = list_kind(SK_FIRE) ? 1 : 0;
        list_fields(sp, s, skin, list_cap);
        MDX_TRY(upload_all_sp());
        if (reorder_every > 0 && s.cells_valid && s.since_reorder >= reorder_every) MDX_TRY(reorder({r}));
        MDX_TRY(run_passes(end - j, SUMS_NONE));
        s.since_reorder += end - j;
        j = end;
        MDX_TRY(read_scalars(r, h));
        const bool sample = (j - 1) >= interval && (j - 1) % interval == 0 && h.f_drifted >= j - 1;
        if (sample || h.f_done) MDX_TRY(potential(U));
        if (sample) push(j - 1, U);   // :209-215
        if (h.f_done) {
            converged = true;
            push(h.f_conv_step + 1, U);   // :170-178
            break;
        }
        if (fp.fp64_switch > 0.0 && prec != MDX_PREC_FP64 && h.f_maxF < fp.fp64_switch) {
            MDX_TRY(set_precision(MDX_PREC_FP64));
        }
    }
    if (!converged && last_pushed != max_steps) {   // :227-233
        MDX_TRY(potential(U));
        push(max_steps, U);
    }
    res->steps = converged ? h.f_conv_step : max_steps;
    res->converged = converged ? 1 : 0;
    res->max_force = h.f_maxF;
    MDX_TRY(potential(U));   // :270
    res->potential_energy = U;
    res->n_energy = ns;
    res->pad = 0;
    set_idle(sp, s);
    s.kind = SK_NONE;
    s.K_valid = s.W_valid = false;
    return upload_sp(r);
}

}  // namespace mdx
