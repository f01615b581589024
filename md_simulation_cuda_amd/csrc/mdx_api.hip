// extern "C" entry points of include/mdx.h.  A single-system handle is a
// batch with one slot; the batch entry points expose many slots advanced by
// the same launches.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "mdx.h"
#include "mdx_batch.hpp"

namespace mdx {
const char *last_error();
int mdx_cell_capacity(const mdx_params &p, double ratio);
}  // namespace mdx

using mdx::Batch;
using mdx::set_error;

struct mdx_sim {
    Batch b;
};

struct mdx_batch {
    Batch b;
};

namespace {

int to_kind(int32_t k) {
    switch (k) {
    case MDX_NVE: return mdx::SK_NVE;
    case MDX_NVT: return mdx::SK_NVT;
    case MDX_NPH: return mdx::SK_NPH;
    case MDX_PISTON: return mdx::SK_PISTON;
    case MDX_IDLE: return mdx::SK_NONE;
    default: return -1;
    }
}

#define CHECK_HANDLE(h) \
    if (!(h)) return set_error(MDX_ERR_INVALID, "null handle")

int timing_enable(Batch &b, int32_t enable) {
    int rc = b.sync();
    if (rc != MDX_OK) return rc;
    for (auto *vec : {&b.force_ev, &b.step_ev}) {
        for (auto &e : *vec) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        vec->clear();
    }
    b.force_ms = b.step_ms = 0.0;
    b.force_n = b.step_n = 0;
    b.timing = enable != 0;
    return MDX_OK;
}

int timing_read(Batch &b, mdx_timing *out) {
    int rc = b.sync();
    if (rc != MDX_OK) return rc;
    auto drain = [](std::vector<std::pair<hipEvent_t, hipEvent_t>> &vec, double &ms, int64_t *cnt) {
        for (auto &e : vec) {
            float t = 0.f;
            if (hipEventElapsedTime(&t, e.first, e.second) == hipSuccess) {
                ms += t;
                if (cnt) ++*cnt;
            }
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        vec.clear();
    };
    drain(b.force_ev, b.force_ms, &b.force_n);
    drain(b.step_ev, b.step_ms, nullptr);
    out->force_ms = b.force_ms;
    out->force_launches = b.force_n;
    out->step_ms = b.step_ms;
    out->steps = b.step_n;
    return MDX_OK;
}

}  // namespace

extern "C" {

const char *mdx_last_error(void) { return mdx::last_error(); }

const char *mdx_version(void) { return "mdx 0.2 (gfx950, batched slots)"; }

int mdx_device_count(int32_t *count) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    if (count) *count = c;
    return MDX_OK;
}

int mdx_cell_capacity(const mdx_params *params, double growth, int32_t *out) {
    if (!params || !out) return set_error(MDX_ERR_INVALID, "null argument");
    if (!(params->Lx > 0.0) || !(params->Ly > 0.0) || !(params->cutoff > 0.0))
        return set_error(MDX_ERR_INVALID, "invalid box or cutoff");
    const int c = mdx::mdx_cell_capacity(*params, growth);
    if (c < 0) return set_error(MDX_ERR_CAPACITY, "cell grid above 2^30 cells");
    *out = c;
    return MDX_OK;
}

int mdx_create(const mdx_params *params, const mdx_particle64 *frame, int64_t n, mdx_sim **out) {
    if (!params || !frame || !out) return set_error(MDX_ERR_INVALID, "null argument");
    *out = nullptr;
    if (n <= 0) return set_error(MDX_ERR_INVALID, "n must be positive");
    if (n > (int64_t{1} << 29) - 256) return set_error(MDX_ERR_CAPACITY, "n above 2^29 - 256");
    if (!(params->Lx > 0.0) || !(params->Ly > 0.0) || !std::isfinite(params->Lx) || !std::isfinite(params->Ly))
        return set_error(MDX_ERR_BOX, "box sides must be positive and finite");
    if (!(params->cutoff > 0.0) || !std::isfinite(params->cutoff))
        return set_error(MDX_ERR_INVALID, "cutoff must be > 0");
    // cells for the largest box the barostats may reach
    const int ccap = mdx::mdx_cell_capacity(*params, std::max(1.0, params->barostat_height_max_ratio));
    if (ccap < 0) return set_error(MDX_ERR_CAPACITY, "cell grid above 2^30 cells");
    auto s = std::make_unique<mdx_sim>();
    int32_t cc = ccap;
    int rc = s->b.init(1, &n, &cc, params->precision, params->device);
    if (rc != MDX_OK) return rc;
    rc = s->b.load(0, *params, frame, n);
    if (rc != MDX_OK) return rc;
    rc = s->b.sync();
    if (rc != MDX_OK) return rc;
    *out = s.release();
    return MDX_OK;
}

int mdx_destroy(mdx_sim *sim) {
    delete sim;
    return MDX_OK;
}

int mdx_set_state(mdx_sim *sim, const mdx_particle64 *frame, int64_t n) {
    CHECK_HANDLE(sim);
    if (!frame) return set_error(MDX_ERR_INVALID, "null frame");
    const int rc = sim->b.set_state(0, frame, n);
    if (rc != MDX_OK) return rc;
    return sim->b.sync();
}

int mdx_step(mdx_sim *sim, int32_t kind, int32_t nsteps, double p_ext, uint32_t flags) {
    CHECK_HANDLE(sim);
    if (nsteps < 0) return set_error(MDX_ERR_INVALID, "nsteps < 0");
    const int k = to_kind(kind);
    if (k < 0 || k == mdx::SK_NONE) return set_error(MDX_ERR_INVALID, "unknown ensemble");
    const int rc = sim->b.set_kind(0, k, p_ext, flags);
    if (rc != MDX_OK) return rc;
    return sim->b.step(nsteps);
}

int mdx_fire(mdx_sim *sim, const mdx_fire_params *fp, double force_tol, int32_t max_steps,
             mdx_fire_result *result, int32_t *energy_steps, double *energy_trace, int32_t cap) {
    CHECK_HANDLE(sim);
    if (!fp || !result) return set_error(MDX_ERR_INVALID, "null argument");
    if (cap > 0 && (!energy_steps || !energy_trace)) return set_error(MDX_ERR_INVALID, "null energy buffers");
    return sim->b.fire(0, *fp, force_tol, max_steps, result, energy_steps, energy_trace, cap);
}

int mdx_observables(mdx_sim *sim, uint32_t mask, mdx_obs *out) {
    CHECK_HANDLE(sim);
    if (!out) return set_error(MDX_ERR_INVALID, "null argument");
    return sim->b.observables(0, mask, out);
}

int mdx_partial_U_lambda(mdx_sim *sim, double eps_lambda, double *out) {
    CHECK_HANDLE(sim);
    if (!out) return set_error(MDX_ERR_INVALID, "null argument");
    return sim->b.partial_U_lambda(0, eps_lambda, out);
}

int mdx_collect(mdx_sim *sim, mdx_particle64 *out, int64_t cap) {
    CHECK_HANDLE(sim);
    if (!out) return set_error(MDX_ERR_INVALID, "null argument");
    return sim->b.collect(0, out, cap);
}

int mdx_compute_forces(mdx_sim *sim, int32_t periodic_y, double *acc_out) {
    CHECK_HANDLE(sim);
    int rc = sim->b.compute_forces({0}, periodic_y ? 1 : 0, false);
    if (rc != MDX_OK) return rc;
    if (acc_out) {
        const int64_t n = sim->b.slots[0].n;
        std::vector<mdx_particle64> tmp(n);
        rc = sim->b.collect(0, tmp.data(), n);
        if (rc != MDX_OK) return rc;
        for (int64_t i = 0; i < n; ++i) {
            acc_out[2 * i] = tmp[i].acc[0];
            acc_out[2 * i + 1] = tmp[i].acc[1];
        }
    }
    return sim->b.sync();
}

int mdx_get_n(const mdx_sim *sim, int64_t *n) {
    CHECK_HANDLE(sim);
    if (!n) return set_error(MDX_ERR_INVALID, "null argument");
    *n = sim->b.slots[0].n;
    return MDX_OK;
}

int mdx_set_precision(mdx_sim *sim, int32_t precision) {
    CHECK_HANDLE(sim);
    return sim->b.set_precision(precision);
}

int mdx_set_dt(mdx_sim *sim, double dt) {
    CHECK_HANDLE(sim);
    if (!std::isfinite(dt)) return set_error(MDX_ERR_INVALID, "dt must be finite");
    sim->b.slots[0].p.dt = dt;
    sim->b.h_sp[0].dt = dt;
    return sim->b.upload_sp(0);
}

int mdx_set_graph(mdx_sim *sim, int32_t enable) {
    CHECK_HANDLE(sim);
    sim->b.use_graph = enable != 0;
    return MDX_OK;
}

int mdx_set_reorder(mdx_sim *sim, int32_t steps) {
    CHECK_HANDLE(sim);
    if (steps < 0) return set_error(MDX_ERR_INVALID, "reorder interval must be >= 0");
    sim->b.reorder_every = steps;
    return MDX_OK;
}

int mdx_sync(mdx_sim *sim) {
    CHECK_HANDLE(sim);
    return sim->b.sync();
}

int mdx_cell_list(mdx_sim *sim, int32_t periodic_y, int32_t *ncx, int32_t *ncy, uint32_t *keys,
                  uint32_t *cell_start, uint32_t *sorted, int64_t cap_n, int64_t cap_cells) {
    CHECK_HANDLE(sim);
    return sim->b.cell_list(0, periodic_y ? 1 : 0, ncx, ncy, keys, cell_start, sorted, cap_n, cap_cells);
}

int mdx_N_profile(mdx_sim *sim, int32_t n_bins, int32_t *counts) {
    CHECK_HANDLE(sim);
    return sim->b.n_profile(0, n_bins, counts);
}

int mdx_pressure_profile(mdx_sim *sim, int32_t n_bins, double *out) {
    CHECK_HANDLE(sim);
    return sim->b.pressure_profile(0, n_bins, out);
}

int mdx_middle_reflect(mdx_sim *sim) {
    CHECK_HANDLE(sim);
    return sim->b.middle_reflect(0);
}

int mdx_N_profile_noalloc(mdx_sim *sim, int32_t n_bins, int32_t *counts) {
    CHECK_HANDLE(sim);
    return sim->b.n_profile_noalloc(0, n_bins, counts);
}

int mdx_timing_enable(mdx_sim *sim, int32_t enable) {
    CHECK_HANDLE(sim);
    return timing_enable(sim->b, enable);
}

int mdx_timing_read(mdx_sim *sim, mdx_timing *out) {
    CHECK_HANDLE(sim);
    if (!out) return set_error(MDX_ERR_INVALID, "null argument");
    return timing_read(sim->b, out);
}

int mdx_get_stream(mdx_sim *sim, void **stream) {
    CHECK_HANDLE(sim);
    if (!stream) return set_error(MDX_ERR_INVALID, "null argument");
    *stream = static_cast<void *>(sim->b.stream);
    return MDX_OK;
}

int mdx_set_stream(mdx_sim *sim, void *stream) {
    CHECK_HANDLE(sim);
    return sim->b.set_stream(static_cast<hipStream_t>(stream));
}

int mdx_time_pair_pass(mdx_sim *sim, int32_t reps, double *avg_ms) {
    CHECK_HANDLE(sim);
    if (!avg_ms) return set_error(MDX_ERR_INVALID, "null argument");
    *avg_ms = 0.0;
    return sim->b.time_pair_pass(reps, avg_ms);
}

/* ---- batches ------------------------------------------------------------ */

int mdx_batch_create(int32_t nslots, const int64_t *n_cap, const int32_t *cell_cap, int32_t precision,
                     int32_t device, mdx_batch **out) {
    if (!n_cap || !cell_cap || !out) return set_error(MDX_ERR_INVALID, "null argument");
    *out = nullptr;
    auto b = std::make_unique<mdx_batch>();
    const int rc = b->b.init(nslots, n_cap, cell_cap, precision, device);
    if (rc != MDX_OK) return rc;
    *out = b.release();
    return MDX_OK;
}

int mdx_batch_destroy(mdx_batch *b) {
    delete b;
    return MDX_OK;
}

int mdx_batch_load(mdx_batch *b, int32_t slot, const mdx_params *params, const mdx_particle64 *frame,
                   int64_t n) {
    CHECK_HANDLE(b);
    if (!params) return set_error(MDX_ERR_INVALID, "null params");
    if (params->precision != b->b.prec)
        return set_error(MDX_ERR_INVALID, "slot precision differs from the batch precision");
    const int rc = b->b.load(slot, *params, frame, n);
    if (rc != MDX_OK) return rc;
    return b->b.sync();
}

int mdx_batch_unload(mdx_batch *b, int32_t slot) {
    CHECK_HANDLE(b);
    return b->b.unload(slot);
}

int mdx_batch_set_kind(mdx_batch *b, int32_t slot, int32_t kind, double p_ext, uint32_t flags) {
    CHECK_HANDLE(b);
    const int k = to_kind(kind);
    if (k < 0) return set_error(MDX_ERR_INVALID, "unknown ensemble");
    return b->b.set_kind(slot, k, p_ext, flags);
}

int mdx_batch_step(mdx_batch *b, int32_t nsteps) {
    CHECK_HANDLE(b);
    return b->b.step(nsteps);
}

int mdx_batch_observables(mdx_batch *b, int32_t slot, uint32_t mask, mdx_obs *out) {
    CHECK_HANDLE(b);
    if (!out) return set_error(MDX_ERR_INVALID, "null argument");
    return b->b.observables(slot, mask, out);
}

int mdx_batch_collect(mdx_batch *b, int32_t slot, mdx_particle64 *out, int64_t cap) {
    CHECK_HANDLE(b);
    if (!out) return set_error(MDX_ERR_INVALID, "null argument");
    return b->b.collect(slot, out, cap);
}

int mdx_batch_fire(mdx_batch *b, int32_t slot, const mdx_fire_params *fp, double force_tol, int32_t max_steps,
                   mdx_fire_result *result, int32_t *energy_steps, double *energy_trace, int32_t cap) {
    CHECK_HANDLE(b);
    if (!fp || !result) return set_error(MDX_ERR_INVALID, "null argument");
    if (cap > 0 && (!energy_steps || !energy_trace)) return set_error(MDX_ERR_INVALID, "null energy buffers");
    return b->b.fire(slot, *fp, force_tol, max_steps, result, energy_steps, energy_trace, cap);
}

int mdx_batch_sync(mdx_batch *b) {
    CHECK_HANDLE(b);
    return b->b.sync();
}

int mdx_batch_set_graph(mdx_batch *b, int32_t enable) {
    CHECK_HANDLE(b);
    b->b.use_graph = enable != 0;
    return MDX_OK;
}

int mdx_batch_set_reorder(mdx_batch *b, int32_t steps) {
    CHECK_HANDLE(b);
    if (steps < 0) return set_error(MDX_ERR_INVALID, "reorder interval must be >= 0");
    b->b.reorder_every = steps;
    return MDX_OK;
}

int mdx_batch_timing_enable(mdx_batch *b, int32_t enable) {
    CHECK_HANDLE(b);
    return timing_enable(b->b, enable);
}

int mdx_batch_timing_read(mdx_batch *b, mdx_timing *out) {
    CHECK_HANDLE(b);
    if (!out) return set_error(MDX_ERR_INVALID, "null argument");
    return timing_read(b->b, out);
}

int mdx_set_tuning(mdx_sim *sim, const mdx_tuning *t) {
    CHECK_HANDLE(sim);
    if (!t) return set_error(MDX_ERR_INVALID, "null argument");
    return sim->b.set_tuning(*t);
}

int mdx_get_tuning(mdx_sim *sim, mdx_tuning *out) {
    CHECK_HANDLE(sim);
    if (!out) return set_error(MDX_ERR_INVALID, "null argument");
    *out = sim->b.tuning();
    return MDX_OK;
}

int mdx_autotune(mdx_sim *sim, int32_t kind, int32_t steps, mdx_tune_report *out) {
    CHECK_HANDLE(sim);
    if (!out) return set_error(MDX_ERR_INVALID, "null argument");
    const int k = to_kind(kind);
    if (k < 0) return set_error(MDX_ERR_INVALID, "unknown ensemble");
    return sim->b.autotune(steps, out, k);
}

int mdx_batch_set_tuning(mdx_batch *b, const mdx_tuning *t) {
    CHECK_HANDLE(b);
    if (!t) return set_error(MDX_ERR_INVALID, "null argument");
    return b->b.set_tuning(*t);
}

int mdx_batch_autotune(mdx_batch *b, int32_t steps, mdx_tune_report *out) {
    CHECK_HANDLE(b);
    if (!out) return set_error(MDX_ERR_INVALID, "null argument");
    return b->b.autotune(steps, out, -1);
}

int mdx_batch_fire_begin(mdx_batch *b, int32_t slot, const mdx_fire_params *fp, double force_tol,
                         int32_t max_steps) {
    CHECK_HANDLE(b);
    if (!fp) return set_error(MDX_ERR_INVALID, "null argument");
    return b->b.fire_begin(slot, *fp, force_tol, max_steps);
}

int mdx_batch_fire_poll(mdx_batch *b, int32_t slot, int32_t *done, mdx_fire_result *result) {
    CHECK_HANDLE(b);
    if (!done || !result) return set_error(MDX_ERR_INVALID, "null argument");
    int d = 0;
    const int rc = b->b.fire_poll(slot, &d, result);
    *done = d;
    return rc;
}

}  // extern "C"
