// Job server: continuous batching of independent MD jobs over device-resident
// slots.  No counterpart in the reference, whose drivers (run/*.cpp) run one
// system per process from start to end.
//
// Three batches ("pools") hold fixed slots in power-of-two size classes: NVE,
// NVT and NPH jobs in a batch of the server's MD precision with neighbour
// lists (rebuilt on demand, so a job's rebuild points depend only on its own
// particles; for NPH the rebuild test follows the barostat's scaling), piston
// jobs in a batch of that precision with the cell list every step, FIRE jobs
// in an FP64-accelerating batch (FIRE needs accelerations beyond FP32 range
// at overlapping starts), with neighbour lists rebuilt on demand when
// fire_skin > 0.  Each tick:
//   1. admits queued jobs in arrival order into the smallest free slot that
//      fits them, while the bytes of the admitted slots stay within the HBM
//      budget (a job that does not fit now does not block later ones);
//   2. advances every running job by the same number of steps q (at most the
//      quantum and at most any MD job's remaining steps); every pool's
//      launches are queued before the host waits on any of them, so the pools
//      overlap on the device;
//   3. records K, U, P of every running MD job (one kinetic and one energy
//      pass per pool), copies the current frame of each job that streams
//      frames (frame_interval f > 0) and whose step count passed a multiple of
//      f in this tick, and retires finished jobs: final observables, frame
//      copied to the host, slot freed.
// Because every reduction is per slot and a slot's neighbour order depends only
// on its own particles, a job gives the same result as the same system run
// alone through mdx_create / mdx_step / mdx_fire.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <vector>

#include "mdx.h"
#include "mdx_batch.hpp"

namespace mdx {
int mdx_cell_capacity(const mdx_params &p, double ratio);
}

using mdx::Batch;
using mdx::set_error;

namespace {

constexpr int kPoolMD = 0, kPoolFire = 1, kPoolList = 2, kPools = 3;

// Device bytes one slot of capacity n particles and c cells occupies (the
// batch's per-particle and per-cell arrays at the given precision, and the
// neighbour-list arrays when it has them).
double slot_bytes(int64_t n, int32_t c, int prec, bool lists) {
    const double a = prec == MDX_PREC_COMP32 ? 8.0 : 16.0;
    const double img = prec == MDX_PREC_FP64 ? 16.0 : 8.0;
    // x, v, a, image, class x2, key, sorted, cell of slot, sorted_old, movers
    double per_particle = 16.0 + 16.0 + a + img + 2.0 + 4.0 * 5.0;
    // 64 list entries, list length, key and position and image at the build
    if (lists) per_particle += 64.0 * 4.0 + 4.0 + 4.0 + 16.0 + 8.0;
    // count, cell_start, cell_start_old, mover offsets
    const double per_cell = 4.0 * 4.0;
    return per_particle * static_cast<double>(n) + per_cell * static_cast<double>(c);
}

struct Job {
    int64_t id = 0;
    mdx_job spec{};
    std::vector<mdx_particle64> frame;   // input until admitted, then the result
    int32_t state = MDX_JOB_QUEUED;
    int pool = -1, slot = -1;
    int64_t n = 0, steps_done = 0;
    double bytes = 0.0;
    double t_arrival = 0.0, t_admit = -1.0, t_done = -1.0;
    mdx_obs last{};
    mdx_fire_result fire{};
    std::vector<mdx_energy_record> records;
    struct Streamed {
        int64_t step;
        std::vector<mdx_particle64> frame;
    };
    std::deque<Streamed> frames;   // streamed frames not yet taken, oldest first
};

struct Pool {
    Batch b;
    int prec = MDX_PREC_COMP32;
    std::vector<int64_t> cap;
    std::vector<int32_t> ccap;
    std::vector<int64_t> job;   // job id in the slot, -1 free
};

}  // namespace

struct mdx_server {
    mdx_server_params p{};
    Pool pools[kPools];
    bool lists = false;        // the NVE / NVT / NPH pool has neighbour lists
    bool fire_lists = false;   // the FIRE pool has neighbour lists
    std::map<int64_t, Job> jobs;
    std::deque<int64_t> queue;
    int64_t next_id = 1;
    double reserved = 0.0, peak_reserved = 0.0;
    double atom_steps = 0.0;
    int64_t ticks = 0, completed = 0;
    int64_t frames_streamed = 0, frames_dropped = 0;
    std::chrono::steady_clock::time_point t0;

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    int init(const mdx_server_params &sp) {
        p = sp;
        t0 = std::chrono::steady_clock::now();
        const int counts[kPools] = {p.md_slots_per_class, p.fire_slots_per_class, p.list_slots_per_class};
        const int precs[kPools] = {p.md_precision, MDX_PREC_ACC64, p.md_precision};
        for (int q = 0; q < kPools; ++q) {
            Pool &pl = pools[q];
            pl.prec = precs[q];
            for (int k = p.min_class_log2; k <= p.max_class_log2; ++k) {
                const int64_t cap = int64_t{1} << k;
                // cells of the largest box a job of this size may reach: a
                // square box at the minimum density, grown by the barostat clamp
                mdx_params bp{};
                bp.Lx = bp.Ly = std::sqrt(static_cast<double>(cap) / p.min_density);
                bp.sigma_aa = bp.sigma_bb = bp.sigma_ab = 1.0;
                bp.cutoff = 2.5;
                bp.cell_scale = 1.0;
                const int c = mdx::mdx_cell_capacity(bp, p.max_box_growth);
                if (c < 0) return set_error(MDX_ERR_CAPACITY, "server: cell capacity overflow");
                for (int s = 0; s < counts[q]; ++s) {
                    pl.cap.push_back(cap);
                    pl.ccap.push_back(c + c / 8 + 64);
                }
            }
            pl.job.assign(pl.cap.size(), -1);
            if (pl.cap.empty()) continue;
            int rc = pl.b.init(static_cast<int>(pl.cap.size()), pl.cap.data(), pl.ccap.data(), pl.prec, p.device);
            if (rc != MDX_OK) return rc;
            if (q == kPoolList && p.list_skin > 0.0 && pl.prec != MDX_PREC_FP64) {
                rc = pl.b.set_tuning(mdx_tuning{1.0, p.list_skin, 1, 0});
                if (rc != MDX_OK) return rc;
                lists = true;
            }
            if (q == kPoolFire && p.fire_skin > 0.0) {   // rebuilt when an iteration needs it
                rc = pl.b.set_tuning(mdx_tuning{1.0, p.fire_skin, 1, 0});
                if (rc != MDX_OK) return rc;
                fire_lists = true;
            }
        }
        return MDX_OK;
    }

    // the pool a job of this kind runs in
    int pool_of(int32_t kind) const {
        if (kind == MDX_FIRE) return kPoolFire;
        if ((kind == MDX_NVE || kind == MDX_NVT || kind == MDX_NPH) && !pools[kPoolList].cap.empty())
            return kPoolList;
        return kPoolMD;
    }

    int submit(const mdx_job &j, const mdx_particle64 *frame, int64_t n, int64_t *id) {
        if (n <= 0) return set_error(MDX_ERR_INVALID, "server: job with no particles");
        if (j.nsteps <= 0) return set_error(MDX_ERR_INVALID, "server: nsteps must be positive");
        const bool fire = j.kind == MDX_FIRE;
        if (!fire && (j.kind < MDX_NVE || j.kind > MDX_PISTON))
            return set_error(MDX_ERR_INVALID, "server: unknown job kind");
        if (j.frame_interval < 0) return set_error(MDX_ERR_INVALID, "server: negative frame_interval");
        const Pool &pl = pools[pool_of(j.kind)];
        if (pl.cap.empty() || n > pl.cap.back())
            return set_error(MDX_ERR_CAPACITY, "server: job larger than the largest slot");
        Job job;
        job.id = next_id++;
        job.spec = j;
        job.n = n;
        job.frame.assign(frame, frame + n);
        job.t_arrival = j.t_arrival >= 0.0 ? j.t_arrival : now();
        *id = job.id;
        queue.push_back(job.id);
        jobs.emplace(job.id, std::move(job));
        return MDX_OK;
    }

    // smallest free slot of the job's pool that holds its particles and cells
    int find_slot(const Job &j, int pool, int32_t need_cells) const {
        const Pool &pl = pools[pool];
        int best = -1;
        for (size_t s = 0; s < pl.cap.size(); ++s) {
            if (pl.job[s] >= 0 || pl.cap[s] < j.n || pl.ccap[s] < need_cells) continue;
            if (best < 0 || pl.cap[s] < pl.cap[best]) best = static_cast<int>(s);
        }
        return best;
    }

    int admit() {
        for (auto it = queue.begin(); it != queue.end();) {
            Job &j = jobs.at(*it);
            const int pool = pool_of(j.spec.kind);
            Pool &pl = pools[pool];
            const mdx_params &jp = j.spec.params;
            const int need = mdx::mdx_cell_capacity(jp, std::max(1.0, jp.barostat_height_max_ratio));
            if (need < 0) return set_error(MDX_ERR_CAPACITY, "server: cell grid above 2^30 cells");
            const int s = find_slot(j, pool, need);
            if (s < 0) {
                ++it;
                continue;
            }
            const double bytes = slot_bytes(pl.cap[s], pl.ccap[s], pl.prec,
                                            (pool == kPoolList && lists) || (pool == kPoolFire && fire_lists));
            if (reserved + bytes > p.hbm_budget_bytes) {
                ++it;
                continue;
            }
            mdx_params lp = jp;
            lp.precision = pl.prec;
            int rc = pl.b.load(s, lp, j.frame.data(), j.n);
            if (rc != MDX_OK) return rc;
            if (j.spec.kind == MDX_FIRE) {
                rc = pl.b.fire_begin(s, j.spec.fire, j.spec.force_tol, j.spec.nsteps);
            } else {
                const int k = j.spec.kind == MDX_NVE   ? mdx::SK_NVE
                              : j.spec.kind == MDX_NVT ? mdx::SK_NVT
                              : j.spec.kind == MDX_NPH ? mdx::SK_NPH
                                                       : mdx::SK_PISTON;
                rc = pl.b.set_kind(s, k, j.spec.p_ext, 0);
            }
            if (rc != MDX_OK) return rc;
            pl.job[s] = j.id;
            j.pool = pool;
            j.slot = s;
            j.bytes = bytes;
            j.state = MDX_JOB_RUNNING;
            j.t_admit = now();
            reserved += bytes;
            peak_reserved = std::max(peak_reserved, reserved);
            std::vector<mdx_particle64>().swap(j.frame);
            it = queue.erase(it);
        }
        return MDX_OK;
    }

    // the job's current frame into its stream (the oldest dropped beyond
    // frame_queue frames)
    int stream_frame(Job &j) {
        Job::Streamed f{j.steps_done, std::vector<mdx_particle64>(static_cast<size_t>(j.n))};
        int rc = pools[j.pool].b.collect(j.slot, f.frame.data(), j.n);
        if (rc != MDX_OK) return rc;
        j.frames.push_back(std::move(f));
        ++frames_streamed;
        while (static_cast<int64_t>(j.frames.size()) > std::max(1, p.frame_queue)) {
            j.frames.pop_front();
            ++frames_dropped;
        }
        return MDX_OK;
    }
    // the job's count passed a multiple of its frame interval in the last q steps
    static bool frame_due(const Job &j, int64_t q) {
        const int64_t f = j.spec.frame_interval;
        return f > 0 && j.steps_done / f > (j.steps_done - q) / f;
    }

    int retire(Job &j) {
        Pool &pl = pools[j.pool];
        j.frame.resize(j.n);
        int rc = pl.b.collect(j.slot, j.frame.data(), j.n);
        if (rc != MDX_OK) return rc;
        rc = pl.b.unload(j.slot);
        if (rc != MDX_OK) return rc;
        pl.job[j.slot] = -1;
        reserved -= j.bytes;
        j.state = MDX_JOB_DONE;
        j.t_done = now();
        ++completed;
        return MDX_OK;
    }

    int tick(int32_t *n_running, int32_t *n_queued, int64_t *done_ids, int32_t cap_done, int32_t *n_done) {
        int32_t nd = 0;
        int rc = admit();
        if (rc != MDX_OK) return rc;
        // step count of this tick
        int64_t q = std::max(1, p.quantum);
        bool any = false;
        for (auto &pl : pools)
            for (size_t s = 0; s < pl.job.size(); ++s) {
                if (pl.job[s] < 0) continue;
                const Job &j = jobs.at(pl.job[s]);
                const int64_t total = j.spec.kind == MDX_FIRE ? j.spec.nsteps + 1 : j.spec.nsteps;
                q = std::min(q, std::max<int64_t>(1, total - j.steps_done));
                any = true;
            }
        if (any) {
            ++ticks;
            // FIRE pool first, then the cell-list pool, then the neighbour-list
            // pool, whose NPH jobs end their step call with a read of the box
            // flags that waits for that pool's work: the pools' launches are
            // all queued before the host waits, so they overlap on the device
            for (int pi : {kPoolFire, kPoolMD, kPoolList}) {
                Pool &pl = pools[pi];
                if (std::none_of(pl.job.begin(), pl.job.end(), [](int64_t x) { return x >= 0; })) continue;
                rc = pl.b.step(static_cast<int>(q));
                if (rc != MDX_OK) return rc;
            }
            for (auto &pl : pools)
                for (size_t s = 0; s < pl.job.size(); ++s)
                    if (pl.job[s] >= 0) {
                        Job &j = jobs.at(pl.job[s]);
                        j.steps_done += q;
                        atom_steps += static_cast<double>(q) * static_cast<double>(j.n);
                    }
            // MD jobs: energies of every running job, then retire the finished
            for (int pi : {kPoolMD, kPoolList}) {
                Pool &md = pools[pi];
                std::vector<int> running;
                for (size_t s = 0; s < md.job.size(); ++s) {
                    if (md.job[s] < 0) continue;
                    const Job &j = jobs.at(md.job[s]);
                    if (p.record_energies || j.steps_done >= j.spec.nsteps) running.push_back(static_cast<int>(s));
                }
                std::vector<int> plain, piston;
                for (int s : running) (md.b.slots[s].piston_active ? piston : plain).push_back(s);
                std::vector<mdx_obs> ob(plain.size());
                rc = md.b.observables_many(plain, MDX_OBS_K | MDX_OBS_U | MDX_OBS_P | MDX_OBS_VIRIAL, ob.data());
                if (rc != MDX_OK) return rc;
                for (size_t q2 = 0; q2 < plain.size(); ++q2) jobs.at(md.job[plain[q2]]).last = ob[q2];
                for (int s : piston) {
                    rc = md.b.observables(s, MDX_OBS_K | MDX_OBS_U | MDX_OBS_P | MDX_OBS_VIRIAL,
                                          &jobs.at(md.job[s]).last);
                    if (rc != MDX_OK) return rc;
                }
                for (int s : running) {
                    Job &j = jobs.at(md.job[s]);
                    if (p.record_energies) j.records.push_back({j.steps_done, j.last.K, j.last.U, j.last.P});
                    if (frame_due(j, q)) {
                        rc = stream_frame(j);
                        if (rc != MDX_OK) return rc;
                    }
                    if (j.steps_done >= j.spec.nsteps) {
                        rc = retire(j);
                        if (rc != MDX_OK) return rc;
                        if (nd < cap_done && done_ids) done_ids[nd] = j.id;
                        ++nd;
                    }
                }
            }
            // FIRE jobs: converged or out of iterations
            Pool &fp = pools[kPoolFire];
            for (size_t s = 0; s < fp.job.size(); ++s) {
                if (fp.job[s] < 0) continue;
                Job &j = jobs.at(fp.job[s]);
                int done = 0;
                rc = fp.b.fire_poll(static_cast<int>(s), &done, &j.fire);
                if (rc != MDX_OK) return rc;
                if (!done && frame_due(j, q)) {
                    rc = stream_frame(j);
                    if (rc != MDX_OK) return rc;
                }
                if (done) {
                    j.last.U = j.fire.potential_energy;
                    rc = retire(j);
                    if (rc != MDX_OK) return rc;
                    if (nd < cap_done && done_ids) done_ids[nd] = j.id;
                    ++nd;
                }
            }
            rc = admit();   // fill the freed slots before the next tick
            if (rc != MDX_OK) return rc;
        }
        int32_t nr = 0;
        for (auto &pl : pools)
            for (int64_t x : pl.job) nr += x >= 0 ? 1 : 0;
        if (n_running) *n_running = nr;
        if (n_queued) *n_queued = static_cast<int32_t>(queue.size());
        if (n_done) *n_done = nd;
        return MDX_OK;
    }
};

extern "C" {

int mdx_server_default_params(mdx_server_params *out) {
    if (!out) return set_error(MDX_ERR_INVALID, "null argument");
    std::memset(out, 0, sizeof(*out));
    out->hbm_budget_bytes = 288e9;
    out->min_class_log2 = 10;
    out->max_class_log2 = 18;
    out->md_slots_per_class = 2;
    out->fire_slots_per_class = 2;
    out->quantum = 100;
    out->record_energies = 1;
    out->md_precision = MDX_PREC_COMP32;
    out->min_density = 0.3;
    out->max_box_growth = 1.4;
    out->device = -1;
    out->list_slots_per_class = 16;
    out->list_skin = 0.5;
    out->frame_queue = 8;
    out->fire_skin = 0.0;
    return MDX_OK;
}

int mdx_server_create(const mdx_server_params *params, mdx_server **out) {
    if (!params || !out) return set_error(MDX_ERR_INVALID, "null argument");
    *out = nullptr;
    const mdx_server_params &p = *params;
    if (p.min_class_log2 < 0 || p.max_class_log2 > 28 || p.min_class_log2 > p.max_class_log2)
        return set_error(MDX_ERR_INVALID, "server: bad size classes");
    if (p.md_slots_per_class < 0 || p.fire_slots_per_class < 0 || p.list_slots_per_class < 0 ||
        p.md_slots_per_class + p.fire_slots_per_class + p.list_slots_per_class == 0)
        return set_error(MDX_ERR_INVALID, "server: no slots");
    if (!(p.list_skin >= 0.0) || !(p.list_skin <= 4.0)) return set_error(MDX_ERR_INVALID, "server: list_skin in [0, 4]");
    if (!(p.min_density > 0.0) || !(p.max_box_growth >= 1.0) || !(p.hbm_budget_bytes > 0.0))
        return set_error(MDX_ERR_INVALID, "server: bad density, growth or budget");
    if (p.md_precision < MDX_PREC_COMP32 || p.md_precision > MDX_PREC_FP64)
        return set_error(MDX_ERR_INVALID, "server: unknown precision");
    auto s = std::make_unique<mdx_server>();
    const int rc = s->init(p);
    if (rc != MDX_OK) return rc;
    *out = s.release();
    return MDX_OK;
}

int mdx_server_destroy(mdx_server *s) {
    delete s;
    return MDX_OK;
}

int mdx_server_now(mdx_server *s, double *t) {
    if (!s || !t) return set_error(MDX_ERR_INVALID, "null argument");
    *t = s->now();
    return MDX_OK;
}

int mdx_server_submit(mdx_server *s, const mdx_job *job, const mdx_particle64 *frame, int64_t n, int64_t *id) {
    if (!s || !job || !frame || !id) return set_error(MDX_ERR_INVALID, "null argument");
    return s->submit(*job, frame, n, id);
}

int mdx_server_tick(mdx_server *s, int32_t *n_running, int32_t *n_queued, int64_t *done_ids, int32_t cap_done,
                    int32_t *n_done) {
    if (!s) return set_error(MDX_ERR_INVALID, "null handle");
    return s->tick(n_running, n_queued, done_ids, cap_done, n_done);
}

int mdx_server_info(mdx_server *s, int64_t id, mdx_job_info *out) {
    if (!s || !out) return set_error(MDX_ERR_INVALID, "null argument");
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return set_error(MDX_ERR_INVALID, "server: unknown job id");
    const Job &j = it->second;
    std::memset(out, 0, sizeof(*out));
    out->id = j.id;
    out->state = j.state;
    out->kind = j.spec.kind;
    out->n = j.n;
    out->steps_done = j.steps_done;
    out->t_arrival = j.t_arrival;
    out->t_admit = j.t_admit;
    out->t_done = j.t_done;
    out->last = j.last;
    out->fire = j.fire;
    out->n_records = static_cast<int32_t>(j.records.size());
    out->n_frames = static_cast<int32_t>(j.frames.size());
    return MDX_OK;
}

int mdx_server_records(mdx_server *s, int64_t id, mdx_energy_record *out, int32_t cap, int32_t *n) {
    if (!s || !n) return set_error(MDX_ERR_INVALID, "null argument");
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return set_error(MDX_ERR_INVALID, "server: unknown job id");
    const auto &rec = it->second.records;
    const int32_t k = static_cast<int32_t>(std::min<size_t>(rec.size(), static_cast<size_t>(std::max(cap, 0))));
    if (out && k > 0) std::memcpy(out, rec.data(), k * sizeof(mdx_energy_record));
    *n = static_cast<int32_t>(rec.size());
    return MDX_OK;
}

int mdx_server_result(mdx_server *s, int64_t id, mdx_particle64 *frame, int64_t cap, int32_t release) {
    if (!s) return set_error(MDX_ERR_INVALID, "null handle");
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return set_error(MDX_ERR_INVALID, "server: unknown job id");
    Job &j = it->second;
    if (j.state != MDX_JOB_DONE) return set_error(MDX_ERR_STATE, "server: job not done");
    if (frame) {
        if (cap < j.n) return set_error(MDX_ERR_CAPACITY, "output capacity below n");
        std::memcpy(frame, j.frame.data(), j.n * sizeof(mdx_particle64));
    }
    if (release) s->jobs.erase(it);
    return MDX_OK;
}

int mdx_server_stats(mdx_server *s, mdx_server_counters *out) {
    if (!s || !out) return set_error(MDX_ERR_INVALID, "null argument");
    std::memset(out, 0, sizeof(*out));
    out->atom_steps = s->atom_steps;
    out->reserved_bytes = s->reserved;
    out->peak_reserved_bytes = s->peak_reserved;
    out->ticks = s->ticks;
    out->completed = s->completed;
    out->queued = static_cast<int64_t>(s->queue.size());
    int64_t slots = 0;
    for (auto &pl : s->pools) slots += static_cast<int64_t>(pl.cap.size());
    out->slots = slots;
    out->frames_streamed = s->frames_streamed;
    out->frames_dropped = s->frames_dropped;
    return MDX_OK;
}

int mdx_server_frame(mdx_server *s, int64_t id, mdx_particle64 *frame, int64_t cap, int64_t *step, int32_t *got) {
    if (!s || !got) return set_error(MDX_ERR_INVALID, "null argument");
    *got = 0;
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return set_error(MDX_ERR_INVALID, "server: unknown job id");
    Job &j = it->second;
    if (!frame) {   // peek: the number waiting
        *got = static_cast<int32_t>(j.frames.size());
        return MDX_OK;
    }
    if (j.frames.empty()) return MDX_OK;
    if (cap < j.n) return set_error(MDX_ERR_CAPACITY, "output capacity below n");
    std::memcpy(frame, j.frames.front().frame.data(), j.n * sizeof(mdx_particle64));
    if (step) *step = j.frames.front().step;
    j.frames.pop_front();
    *got = 1;
    return MDX_OK;
}

}  // extern "C"
