/*
 * mdx -- C ABI of the MI355X-native 2D binary Lennard-Jones MD step.
 *
 * The reference (Xiaode2333/MD_Simulation_CUDA) exposes this path only as C++
 * classes that launch CUDA kernels on thrust vectors; it has no FFI.  These
 * entry points are the plain-C boundary those classes are rebuilt on: opaque
 * handles, POD parameters, int status codes with a thread-local message from
 * mdx_last_error().  The header-compatible C++ classes (md_env.hpp,
 * md_energy_minimizer.hpp) call only what is declared here.
 *
 * Threading: calls on distinct handles may run concurrently from different
 * host threads; calls on one handle must not overlap.  Device memory belongs to
 * the handle.  Every call that returns data to the host synchronises the
 * handle's stream; stepping calls may return before the GPU finishes.
 */
#ifndef MDX_H
#define MDX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define MDX_OK              0
#define MDX_ERR_INVALID     1   /* bad argument or configuration            */
#define MDX_ERR_HIP         2   /* HIP runtime error                         */
#define MDX_ERR_CAPACITY    3   /* size above a stated maximum              */
#define MDX_ERR_BOX         4   /* invalid box (where the reference aborts) */
#define MDX_ERR_NONFINITE   5   /* non-finite state detected               */
#define MDX_ERR_STATE       6   /* call not valid in the current state     */
#define MDX_ERR_NODEVICE    7   /* no usable GPU                            */

/* ---- data types -------------------------------------------------------- */

/* Byte-identical to the reference's Particle (include/md_particle.hpp:31-38):
 * double2 pos @0, vel @16, acc @32 (acc = F/m), int type @48, 4 pad bytes,
 * double theta @56; 64 bytes.  This is also the PFRAME v2 frame payload. */
typedef struct mdx_particle64 {
    double pos[2];
    double vel[2];
    double acc[2];
    int32_t type;
    int32_t pad;
    double theta;
} mdx_particle64;

typedef enum mdx_ensemble {
    MDX_NVE = 0,      /* step_single_NVE            src/md_env.cu:1576-1637 */
    MDX_NVT = 1,      /* step_single_nose_hoover    src/md_env.cu:1639-1728 */
    MDX_NPH = 2,      /* step_single_NPH            src/md_env.cu:1795-1918 */
    MDX_PISTON = 3,   /* step_single_NPH_piston     src/md_env.cu:1920-2023 */
    MDX_IDLE = -1     /* batch slot that does not advance              */
} mdx_ensemble;

/* Arithmetic of the pair pass. */
typedef enum mdx_precision {
    MDX_PREC_COMP32 = 0,   /* FP32 pair, FP32 two-sum (compensated) accumulation (default) */
    MDX_PREC_ACC64 = 1,    /* FP32 pair, FP64 accumulation (the reference's form)         */
    MDX_PREC_FP64 = 2      /* FP64 pair, FP64 accumulation                                  */
} mdx_precision;

/* mdx_step flags */
#define MDX_STEP_MIDDLE_WRAP 1u   /* NVT only: middle_reflect_LG after the drift */

/* Physical and numerical parameters (the subset of MDConfig the step reads,
 * include/md_config.hpp:14-75). */
typedef struct mdx_params {
    double Lx, Ly;                     /* box_w_global, box_h_global */
    double sigma_aa, sigma_bb, sigma_ab;
    double eps_aa, eps_bb, eps_ab;
    double mass_a, mass_b;             /* type 0 uses mass_a, every other type mass_b */
    double cutoff;                     /* dimensionless: rc_ij = cutoff * sigma_ij */
    double dt;
    double T_target;
    double Q;                          /* Nose-Hoover mass */
    double barostat_mass;              /* NPH area-barostat / piston mass W */
    double k_piston;                   /* piston wall stiffness */
    double barostat_area_rate_init;
    double barostat_height_min_ratio;
    double barostat_height_max_ratio;
    double devide_p;                   /* slab fraction for middle_reflect_LG */
    double cell_scale;                 /* cell side >= cell_scale * cutoff * sigma_max (>= 1; 0 -> 1) */
    int32_t precision;                 /* mdx_precision */
    int32_t device;                    /* HIP device ordinal (-1: current) */
} mdx_params;

/* Observables.  Filled by mdx_observables according to mask. */
#define MDX_OBS_K        1u   /* cal_total_K                    */
#define MDX_OBS_U        2u   /* cal_total_U (+wall energy in piston mode) */
#define MDX_OBS_P        4u   /* cal_instant_pressure (updates last_P)     */
#define MDX_OBS_VIRIAL   8u   /* 1/2 sum_ordered r.F at the current positions */

typedef struct mdx_obs {
    double K, U, P, W;
    double xi;          /* Nose-Hoover variable                 */
    double p_A;         /* NPH area momentum                    */
    double area_rate;   /* get_nph_area_rate(): p_A / W         */
    double v_p;         /* get_nph_piston_velocity()            */
    double F_p;         /* get_nph_piston_force()               */
    double last_P;      /* get_last_instant_pressure()          */
    double Lx, Ly;      /* get_Lx(), get_Ly()                   */
    double t;           /* simulated time                       */
    int32_t piston_active;
    int32_t pad;
    double eta;         /* integral of xi dt over the NVT steps (no counterpart in the
                           reference): K + U + Q xi^2 / 2 + (2N - 2) T0 eta is the
                           Nose-Hoover conserved energy */
} mdx_obs;

/* FIRE parameters (include/md_energy_minimizer.hpp:11-25) plus the documented
 * overlap policy of this implementation. */
typedef struct mdx_fire_params {
    double dt_init, dt_max, alpha0, f_inc, f_dec;
    int32_t N_min;
    int32_t energy_interval;   /* potential sampled every this many steps (reference: 1000) */
    double max_step;           /* cap on |dt v| per step, 0 = uncapped (reference behaviour) */
    double fp64_switch;        /* switch the pair pass to FP64 once max|a| < this (0 = never) */
} mdx_fire_params;

typedef struct mdx_fire_result {
    int32_t steps;
    int32_t converged;
    double max_force;
    double potential_energy;
    int32_t n_energy;          /* samples produced (may exceed the caller's cap) */
    int32_t pad;
} mdx_fire_result;

typedef struct mdx_sim mdx_sim;

/* ---- lifecycle --------------------------------------------------------- */

/* Replaces MDSimulation(cfg, comm) (src/md_env.cu:21-79) minus init_particles:
 * uploads the frame, builds the cell list and computes the initial
 * accelerations (cal_forces, src/md_env.cu:1511-1542). */
int mdx_create(const mdx_params *params, const mdx_particle64 *frame, int64_t n,
               mdx_sim **out);
int mdx_destroy(mdx_sim *sim);

/* Replaces the resume path (src/md_env.cu:81-164) and
 * EnergyMinimizer::initialize_from_frame: replaces positions, velocities and
 * types (n must equal the handle's n), rewraps, recomputes accelerations. */
int mdx_set_state(mdx_sim *sim, const mdx_particle64 *frame, int64_t n);

/* ---- stepping ---------------------------------------------------------- */

/* nsteps steps of one ensemble (mdx_ensemble); p_ext is the external
 * pressure for MDX_NPH / MDX_PISTON.  With rebuild windows (mdx_tuning) a call
 * continues the window the previous call of the same kind left open, so the
 * split into calls does not change the trajectory, and a call does not wait
 * for the GPU; a window violation is then reported by the call itself when it
 * ran at least 64 steps, else by the next mdx_sync, mdx_observables or
 * mdx_collect.  Passes at the current positions (observables, forces) and
 * changes of kind, tuning or state start a new window. */
int mdx_step(mdx_sim *sim, int32_t kind, int32_t nsteps, double p_ext, uint32_t flags);

/* MDSimulation::middle_reflect_LG (include/md_env.hpp:78, src/md_env.cu:
 * 1457-1508): y wrapped, x reflected into the central slab of width
 * devide_p * Lx + 2 sigma_max around Lx / 2; no force pass. */
int mdx_middle_reflect(mdx_sim *sim);

/* FIRE minimisation of the handle's current frame
 * (EnergyMinimizer::minimize_frame, src/md_energy_minimizer.cu:122-272).
 * energy_steps/energy_trace (may be NULL when cap == 0) receive up to cap
 * (step, U) samples.  The minimised frame stays in the handle. */
int mdx_fire(mdx_sim *sim, const mdx_fire_params *fp, double force_tol, int32_t max_steps,
             mdx_fire_result *result, int32_t *energy_steps, double *energy_trace,
             int32_t cap);

/* ---- observables and frames -------------------------------------------- */

/* Getters of src/md_env.cu:2141-2406. */
int mdx_observables(mdx_sim *sim, uint32_t mask, mdx_obs *out);
/* cal_partial_U_lambda(epsilon_lambda), src/md_env.cu:2362-2406. */
int mdx_partial_U_lambda(mdx_sim *sim, double eps_lambda, double *out);

/* Frame in original particle order (sample_collect + get_host_particles,
 * src/md_env.cu:2502-2514).  cap >= n required. */
int mdx_collect(mdx_sim *sim, mdx_particle64 *out, int64_t cap);

/* Recompute accelerations at the current positions without moving anything
 * (cal_forces / cal_forces_piston LJ part); periodic_y selects the y boundary.
 * When acc_out != NULL it receives 2n doubles (ax, ay per particle). */
int mdx_compute_forces(mdx_sim *sim, int32_t periodic_y, double *acc_out);

int mdx_get_n(const mdx_sim *sim, int64_t *n);
/* Replay each step from a captured hipGraph (default on). */
int mdx_set_graph(mdx_sim *sim, int32_t enable);
/* (new) Steps between two reorders of the particle storage into cell order,
 * which keeps the kernels' particle accesses coalesced as particles diffuse
 * away from their load order (default 1000; 0: never).  Results do not
 * depend on it: the cell order sorts by original index and every reduction
 * runs in cell order. */
int mdx_set_reorder(mdx_sim *sim, int32_t steps);
int mdx_set_precision(mdx_sim *sim, int32_t precision);
int mdx_set_dt(mdx_sim *sim, double dt);
int mdx_sync(mdx_sim *sim);

/* Canonical cell list of the current positions (for parity checks):
 * keys[n], cell_start[ncell+1], sorted[n] (original indices, ascending within
 * a cell).  Any array may be NULL; ncx/ncy always written. */
/* Particle counts in n_bins slabs of the box along x, [0, Lx] with the current
 * box: counts[0, n_bins) type 0, counts[n_bins, 2 n_bins) every other type;
 * bin = int(x / Lx * n_bins) clamped to [0, n_bins).  Replaces get_N_profile
 * (src/md_env.cu:3920-3988, kernel src/md_cuda_common.cu:827-877) on one
 * domain: the reference with one rank.  n_bins in [1, 8192]; counts holds
 * 2 n_bins values. */
int mdx_N_profile(mdx_sim *sim, int32_t n_bins, int32_t *counts);
/* Pressure-tensor profile in n_bins slabs of the box along y: out[0, n) Pxx,
 * out[n, 2n) Pyy, out[2n, 3n) Pxy (n = n_bins), each slab's sum over its
 * particles of m v_a v_b and, over the ordered pairs inside the cutoff whose
 * midpoint y_i - d_y / 2 falls in it, of 1/2 d_a F_b, divided by the slab's
 * area Lx Ly / n.  Pair forces in FP64 whatever the precision; sums in fixed
 * point (2^-32), so the result does not depend on summation order.  Replaces
 * get_pressure_profile (src/md_env.cu:3821-3918, kernel
 * src/md_cuda_common.cu:879-1030) on one domain.  y is periodic unless the
 * piston walls are active (the reference always wraps y here).  Rebuilds the
 * cell list: the next step starts a new rebuild window.  n_bins in [1, 8192]. */
int mdx_pressure_profile(mdx_sim *sim, int32_t n_bins, double *out);

/* No counterpart in the reference.  mdx_N_profile for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving or
 * sampling loops, harnesses that forbid hipMalloc): the same counts, kept in a
 * fixed device array of 2 x 8192 int32 counters per GPU instead of a buffer
 * allocated per call, so the call allocates nothing.  n_bins in [1, 8192]
 * (MDX_ERR_CAPACITY beyond).  Every call on one GPU shares that array, so
 * calls that overlap on the same GPU (from several host threads) race; the
 * call synchronises the handle's stream before returning. */
int mdx_N_profile_noalloc(mdx_sim *sim, int32_t n_bins, int32_t *counts);

int mdx_cell_list(mdx_sim *sim, int32_t periodic_y, int32_t *ncx, int32_t *ncy,
                  uint32_t *keys, uint32_t *cell_start, uint32_t *sorted,
                  int64_t cap_n, int64_t cap_cells);

/* ---- kernel timing (HIP events on the handle's stream) ------------------ */
typedef struct mdx_timing {
    double force_ms;        /* summed duration of pair-pass launches timed */
    int64_t force_launches;
    double step_ms;         /* summed duration of whole timed steps        */
    int64_t steps;
} mdx_timing;

/* enable != 0 starts recording (and clears counters); 0 stops. */
int mdx_timing_enable(mdx_sim *sim, int32_t enable);
int mdx_timing_read(mdx_sim *sim, mdx_timing *out);

/* ---- streams ------------------------------------------------------------ */
/* The reference's methods synchronise after every launch; here a handle
 * issues its work on one HIP stream and returns early.  mdx_get_stream gives
 * that stream (a hipStream_t) so a caller can order its own work with the
 * library's (an event recorded on it, hipStreamWaitEvent).  mdx_set_stream
 * makes the handle issue everything after the call on the caller's stream
 * (a hipStream_t of the handle's device, not the null stream; the caller keeps
 * it alive until the handle is destroyed or set back); null returns to the
 * handle's own stream.  Work already issued completes before it returns. */
int mdx_get_stream(mdx_sim *sim, void **stream);
int mdx_set_stream(mdx_sim *sim, void *stream);

/* Measurement support, no counterpart in the reference: the average duration
 * (ms) of the pair pass a step of the current tuning runs (the lean-step
 * variant inside rebuild windows), from reps (1..4096) launches captured back
 * to back in one graph and timed with HIP events on the handle's stream, so
 * no per-launch overhead is included.  The launches run on the current state,
 * which is saved before and restored after (unchanged on return).  Allocates
 * a temporary copy of the velocities, accelerations and scalars; synchronises
 * the handle's stream. */
int mdx_time_pair_pass(mdx_sim *sim, int32_t reps, double *avg_ms);



/* ---- tuning of the pair pass ---------------------------------------------
 * No counterpart in the reference (one fixed launch shape, no cell list).
 * cell_scale: cell side >= cell_scale * cutoff * sigma_max (>= 1).
 * lanes_per_particle: 1, 2 or 4 lanes share one particle's neighbour cells
 * (64, 32 or 16 particles per wavefront); the FP64 form always uses 1.
 * skin: with skin > 0, NVE, NVT and NPH slots and FIRE runs in the FP32 forms
 * keep a list of the partners within cutoff + skin and rebuild it (and the
 * cell list, with cells of at least cutoff + skin) at the step (iteration)
 * where some particle has moved more than skin / 2 since the last build; the
 * pairs evaluated are exactly those within the cutoff, as without the list.
 * FIRE leaves list mode where it switches to FP64 pairs.  NPH: the barostat scales
 * every position by the same factor s between builds, so a particle counts as
 * moved by its distance from its scaled build position, against
 * (s (cutoff + skin) - cutoff) / 2, and a step that changes the cell grid's
 * dimensions rebuilds.  skin = 0: cell list every step.
 * window: with window W > 1 and every stepping slot an NVE or NVT slot in list
 * mode (boxes of at
 * least 5 cells per side), steps run in windows of W: the first (head) step
 * rebuilds a slot's lists when some particle, at 1.5 x its current speed,
 * could move beyond skin / 2 before the window ends; the other W - 1 steps run
 * three kernels (box, a fused kick/drift/image pass, pair).  A lean step in
 * which a slot has a particle beyond skin / 2 anyway (or no usable list)
 * evaluates that slot's forces over the 5 x 5 cells of its last cell list,
 * which holds every pair within the cutoff while no particle has moved more
 * than (2 w - cutoff) / 2 since that build.  If a particle goes beyond that,
 * the run still completes its steps and then returns MDX_ERR_STATE ("use a
 * shorter window"): such a step's forces may have missed pairs.  The run does
 * not synchronise with the host until its end.
 * Rebuild points depend only on W and the slot's own particles.
 * Changing the tuning recomputes the accelerations; FP32 results then follow
 * the new summation order (same tolerance, bit-identical to a fresh run with
 * the same tuning). */
typedef struct mdx_tuning {
    double cell_scale;
    double skin;                 /* > 0: neighbour-list mode (FP32 forms: NVE/NVT/NPH, FIRE) */
    int32_t lanes_per_particle;
    int32_t window;              /* list mode: steps per rebuild window (0: rebuild on demand) */
} mdx_tuning;

#define MDX_TUNE_MAX 32
typedef struct mdx_tune_report {
    mdx_tuning best;
    double best_ms_per_step;
    int32_t n_tried;
    int32_t pad;
    mdx_tuning tried[MDX_TUNE_MAX];
    double ms_per_step[MDX_TUNE_MAX];   /* measured with HIP events over `steps` steps */
} mdx_tune_report;

int mdx_set_tuning(mdx_sim *sim, const mdx_tuning *t);
int mdx_get_tuning(mdx_sim *sim, mdx_tuning *out);
/* Autotuner: on a copy of the system, times `steps` steps of ensemble `kind`
 * for every candidate (without lists: cell_scale 1.0, 1.1 x lanes 1, 2, 4;
 * with lists: skin 0.3, 0.5, 0.65, 0.8 x window 0, 8, 12, 16, 20 with one lane, and
 * skin 0.5, 0.8 with two lanes and on-demand rebuilds; FP64 form: cell
 * list, one lane), then applies the fastest to the handle.  The handle's
 * state is not advanced. */
int mdx_autotune(mdx_sim *sim, int32_t kind, int32_t steps, mdx_tune_report *out);

/* ---- batches of independent systems ---------------------------------------
 * No counterpart in the reference, which runs each sweep point (temperature,
 * epsilon_AB, lambda) as a separate job.  A batch holds nslots slots; slot r
 * holds up to n_cap[r] particles and cell_cap[r] cells (mdx_cell_capacity
 * gives the cells a box needs).  Every mdx_batch_step launch advances all
 * slots whose kind is not MDX_IDLE; each slot's reductions depend only on its
 * own particles, so a system gives the same result alone and in any mix.  All
 * slots share one precision. */
typedef struct mdx_batch mdx_batch;

/* Cells of the canonical grid for the params' box scaled by growth (>= 1). */
int mdx_cell_capacity(const mdx_params *params, double growth, int32_t *out);

int mdx_batch_create(int32_t nslots, const int64_t *n_cap, const int32_t *cell_cap, int32_t precision,
                     int32_t device, mdx_batch **out);
int mdx_batch_destroy(mdx_batch *b);
/* Admit a system into a slot: upload, cell list, initial accelerations. */
int mdx_batch_load(mdx_batch *b, int32_t slot, const mdx_params *params, const mdx_particle64 *frame,
                   int64_t n);
int mdx_batch_unload(mdx_batch *b, int32_t slot);
/* What the slot does on the next mdx_batch_step calls (mdx_ensemble). */
int mdx_batch_set_kind(mdx_batch *b, int32_t slot, int32_t kind, double p_ext, uint32_t flags);
int mdx_batch_step(mdx_batch *b, int32_t nsteps);
int mdx_batch_observables(mdx_batch *b, int32_t slot, uint32_t mask, mdx_obs *out);
int mdx_batch_collect(mdx_batch *b, int32_t slot, mdx_particle64 *out, int64_t cap);
int mdx_batch_fire(mdx_batch *b, int32_t slot, const mdx_fire_params *fp, double force_tol,
                   int32_t max_steps, mdx_fire_result *result, int32_t *energy_steps,
                   double *energy_trace, int32_t cap);
int mdx_batch_sync(mdx_batch *b);
int mdx_batch_set_graph(mdx_batch *b, int32_t enable);
int mdx_batch_set_reorder(mdx_batch *b, int32_t steps);   /* as mdx_set_reorder */
int mdx_batch_timing_enable(mdx_batch *b, int32_t enable);
int mdx_batch_timing_read(mdx_batch *b, mdx_timing *out);
/* FIRE as a slot kind: after mdx_batch_fire_begin the slot takes one FIRE
 * iteration per mdx_batch_step step (EnergyMinimizer::minimize_frame,
 * src/md_energy_minimizer.cu:122-272, without the energy trace) while the
 * other slots do whatever their kind says; max_steps iterations take
 * max_steps + 1 steps.  The batch must be ACC64 or FP64.  mdx_batch_fire_poll
 * sets *done once the slot converged or ran out of iterations, and then also
 * fills result->potential_energy and returns the slot to MDX_IDLE. */
int mdx_batch_fire_begin(mdx_batch *b, int32_t slot, const mdx_fire_params *fp, double force_tol,
                         int32_t max_steps);
int mdx_batch_fire_poll(mdx_batch *b, int32_t slot, int32_t *done, mdx_fire_result *result);
/* Tuning of every loaded slot (cell_scale) and of the batch's pair launches;
 * mdx_batch_autotune times the candidates on a copy of the batch with each
 * slot's current kind (idle slots as NVE) and applies the fastest. */
int mdx_batch_set_tuning(mdx_batch *b, const mdx_tuning *t);
int mdx_batch_autotune(mdx_batch *b, int32_t steps, mdx_tune_report *out);

/* ---- job server: continuous batching of MD jobs -----------------------------
 * No counterpart in the reference, whose drivers (the run/ programs) run one system per
 * process.  Jobs (NVE, NVT, NPH, piston or FIRE, any N up to the largest slot)
 * are submitted at any time; each mdx_server_tick admits queued jobs into free
 * device-resident slots (smallest power-of-two size class that fits, bytes of
 * admitted slots within hbm_budget_bytes, arrival order without head-of-line
 * blocking), advances every running job by up to `quantum` steps in the same
 * launches, records K/U/P of every running MD job, and retires finished jobs
 * (final frame copied to the host, slot freed).  NVE, NVT and NPH jobs run in
 * a batch of md_precision with neighbour lists (skin list_skin, rebuilt on
 * demand), piston jobs in a batch of md_precision with the cell list every
 * step, FIRE jobs in an ACC64 batch (with neighbour lists of skin fire_skin,
 * rebuilt on demand, when fire_skin > 0); the batches' launches overlap on the
 * device.  A job's
 * result is the same as the same system run alone through mdx_create +
 * mdx_step (NVE / NVT / NPH with mdx_set_tuning {1.0, list_skin, 1, 0} in the
 * FP32 forms) or mdx_fire with fp64_switch = 0 (with mdx_set_tuning {1.0,
 * fire_skin, 1, 0}). */
#define MDX_FIRE 4            /* mdx_job.kind for a FIRE minimisation */
#define MDX_JOB_QUEUED 0
#define MDX_JOB_RUNNING 1
#define MDX_JOB_DONE 2

typedef struct mdx_server_params {
    double hbm_budget_bytes;      /* admission budget for resident slots (default 288e9) */
    double min_density;           /* sizes the slots' cell capacity (default 0.3)        */
    double max_box_growth;        /* box growth the slots' cells allow (default 1.4)     */
    int32_t min_class_log2;       /* smallest slot: 2^min_class_log2 particles (10)      */
    int32_t max_class_log2;       /* largest slot (18)                                   */
    int32_t md_slots_per_class;   /* cell-list MD slots (piston jobs) per size class (2) */
    int32_t fire_slots_per_class; /* FIRE slots per size class (2)                       */
    int32_t quantum;              /* most steps per tick (100)                           */
    int32_t record_energies;      /* 1: K/U/P of every running MD job every tick         */
    int32_t md_precision;         /* mdx_precision of the MD batches (COMP32)            */
    int32_t device;               /* -1: current                                         */
    int32_t list_slots_per_class; /* NVE / NVT / NPH slots per size class, in a batch
                                     with neighbour lists (16; 0: they use the MD slots) */
    int32_t frame_queue;          /* streamed frames a job holds until taken (8); a
                                     further frame drops the oldest                      */
    double list_skin;             /* neighbour-list skin of those slots (0.5)           */
    double fire_skin;             /* > 0: neighbour lists for the FIRE slots, rebuilt when
                                     an iteration needs it (0: cell list every
                                     iteration, the default: fast-moving FIRE jobs rebuild
                                     nearly every iteration)                             */
} mdx_server_params;

typedef struct mdx_job {
    mdx_params params;            /* box and interactions; params.precision is ignored   */
    int32_t kind;                 /* MDX_NVE, MDX_NVT, MDX_NPH, MDX_PISTON or MDX_FIRE   */
    int32_t nsteps;               /* MD steps, or FIRE max iterations                    */
    double p_ext;                 /* NPH / piston target pressure                        */
    double force_tol;             /* FIRE: converged when max|a| < force_tol             */
    mdx_fire_params fire;         /* FIRE parameters (energy_interval, fp64_switch unused) */
    double t_arrival;             /* server clock seconds (mdx_server_now); < 0: now     */
    int32_t frame_interval;       /* > 0: stream a frame about every frame_interval MD
                                     steps or FIRE iterations (mdx_server_frame)         */
    int32_t pad;
} mdx_job;

typedef struct mdx_energy_record {
    int64_t step;
    double K, U, P;
} mdx_energy_record;

typedef struct mdx_job_info {
    int64_t id;
    int32_t state;                /* MDX_JOB_* */
    int32_t kind;
    int64_t n, steps_done;
    double t_arrival, t_admit, t_done;   /* server clock seconds; -1 if not yet */
    mdx_obs last;                 /* latest K/U/P (MD) or final PE in U (FIRE)         */
    mdx_fire_result fire;         /* FIRE jobs */
    int32_t n_records;
    int32_t n_frames;             /* streamed frames waiting in mdx_server_frame         */
} mdx_job_info;

typedef struct mdx_server_stats {
    double atom_steps;            /* sum over ticks of steps x N of every running job   */
    double reserved_bytes, peak_reserved_bytes;
    int64_t ticks, completed, queued, slots;
    int64_t frames_streamed, frames_dropped;
} mdx_server_counters;

typedef struct mdx_server mdx_server;

int mdx_server_default_params(mdx_server_params *out);
int mdx_server_create(const mdx_server_params *params, mdx_server **out);
int mdx_server_destroy(mdx_server *s);
int mdx_server_now(mdx_server *s, double *seconds);
int mdx_server_submit(mdx_server *s, const mdx_job *job, const mdx_particle64 *frame, int64_t n,
                      int64_t *job_id);
/* One scheduling round; ids of the jobs it finished go to done_ids[0..cap_done). */
int mdx_server_tick(mdx_server *s, int32_t *n_running, int32_t *n_queued, int64_t *done_ids,
                    int32_t cap_done, int32_t *n_done);
int mdx_server_info(mdx_server *s, int64_t job_id, mdx_job_info *out);
/* Energy records of a job (*n receives the total, at most cap are copied). */
int mdx_server_records(mdx_server *s, int64_t job_id, mdx_energy_record *out, int32_t cap, int32_t *n);
/* Final frame of a finished job (original particle order); release != 0 forgets the job. */
int mdx_server_result(mdx_server *s, int64_t job_id, mdx_particle64 *frame, int64_t cap, int32_t release);
int mdx_server_stats(mdx_server *s, mdx_server_counters *out);
/* Streamed frames: the oldest frame a job with frame_interval > 0 holds
 * (original particle order, as mdx_collect gives it), taken out of its queue.
 * *step receives the job's step (FIRE: iteration) count at the frame, *got 1,
 * or *got 0 when none is waiting.  A frame is taken at the end of each tick in
 * which the job's count reaches or passes a multiple of frame_interval (one
 * frame for several multiples; ticks are not shortened for it), the last MD
 * step included; a job holds at most frame_queue frames.  frame NULL: *got
 * receives the number of frames waiting and none is taken. */
int mdx_server_frame(mdx_server *s, int64_t job_id, mdx_particle64 *frame, int64_t cap, int64_t *step,
                     int32_t *got);

/* ---- initial frames (host only, no GPU) --------------------------------- */

/* Inputs of MDSimulation::init_particles (src/md_env.cu:1276-1445). */
typedef struct mdx_init_params {
    int64_t n_particles_global;
    int64_t n_particles_type0;
    double box_w, box_h;
    double T_init;
    double mass_a, mass_b;
    double devide_p;
} mdx_init_params;

/* The reference constructor's initial frame: phase-separated square grids,
 * std::mt19937(12345) jitter, Maxwell-Boltzmann velocities with the centre of
 * mass removed and K rescaled to (2N-2) T_init / 2, random theta.  Built with
 * libstdc++'s distributions, as the reference is. */
int mdx_init_particles(const mdx_init_params *ip, mdx_particle64 *out);

/* Seeded synthetic frames (std::mt19937_64, own uniform and Box-Muller):
 *   lattice: square box of side sqrt(n/density); ceil(sqrt(n))^2 grid, first n
 *            sites row-major, uniform jitter +-jitter; n_type0 particles of
 *            type 0 chosen by Fisher-Yates; Maxwell-Boltzmann velocities at
 *            T (zero momentum, K = (2n-2) T / 2; T <= 0 gives v = 0).
 *   uniform: the same but positions uniform in the box.
 * box_out receives (Lx, Ly). */
int mdx_gen_lattice(int64_t n, int64_t n_type0, double density, double jitter, double T,
                    double mass_a, double mass_b, uint64_t seed, mdx_particle64 *out,
                    double *box_out);
int mdx_gen_uniform(int64_t n, int64_t n_type0, double density, double T, double mass_a,
                    double mass_b, uint64_t seed, mdx_particle64 *out, double *box_out);

/* ---- misc ---------------------------------------------------------------- */
const char *mdx_last_error(void);
const char *mdx_version(void);
int mdx_device_count(int32_t *count);

#ifdef __cplusplus
}
#endif

#endif /* MDX_H */
