/*
 * FP64 CPU oracle for the 2D binary Lennard-Jones MD step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C++ restatement of the
 * reference's force, observable, integrator and FIRE logic, used as the
 * checker for the HIP library.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as the checker or as the
 * timed CPU baseline -- never as the thing measured or shipped.  The product
 * (libmdx) does not link or call it.
 *
 * Reference semantics restated (file:line relative to the reference repo):
 *   pair force / type selection   src/md_cuda_common.cu:151-264
 *   shifted pair energy           src/md_cuda_common.cu:703-825
 *   kinetic energy                src/md_cuda_common.cu:554-589
 *   scalar pressure (virial)      src/md_cuda_common.cu:591-701, src/md_env.cu:2205-2266
 *   dU/dlambda                    src/md_cuda_common.cu:266-377, src/md_env.cu:2362-2406
 *   VV / NH / NPH / piston kernels src/md_cuda_common.cu:381-552
 *   step sequencing               src/md_env.cu:1576-2023
 *   FIRE                          src/md_energy_minimizer.cu:24-71,122-272,442-475
 *   periodic wrap                 include/md_cuda_common.hpp:20-29
 *
 * What pins each function is listed with the tests (tests/test_oracle.py,
 * tests/test_oracle_pins.py).  FIRE: the fixed point of the reference's
 * two-particle case, the max_steps / immediate-convergence semantics, and the
 * whole dt / alpha / streak schedule, step for step against an independent
 * numpy FIRE written from Bitzek et al. (PRL 97, 170201, 2006) in the
 * reference's step order, on runs that visit every branch of the schedule.
 *
 * Deliberate, documented differences from the reference:
 *   - one periodic domain: every pair is counted once per ordered pair (the
 *     reference's rank_size==1 self-halo double count is not reproduced);
 *   - all pair arithmetic is FP64 (the reference evaluates pairs in FP32);
 *   - FIRE may cap the per-step displacement (max_step > 0); 0 keeps the
 *     reference behaviour.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 64-byte particle record, byte-identical to the reference's Particle
 * (include/md_particle.hpp:31-38): pos@0, vel@16, acc@32, type@48, theta@56. */
typedef struct orc_particle {
    double px, py;
    double vx, vy;
    double ax, ay;
    int32_t type;
    int32_t pad;
    double theta;
} orc_particle;

typedef struct orc_params {
    double Lx, Ly;
    double sigma_aa, sigma_bb, sigma_ab;
    double eps_aa, eps_bb, eps_ab;
    double mass_a, mass_b;
    double cutoff;
    double dt;
    double T_target;
    double Q;
    double barostat_mass;
    double k_piston;
    double barostat_area_rate_init;
    double barostat_height_min_ratio;
    double barostat_height_max_ratio;
    double devide_p;
    double cell_scale;     /* canonical cell list: cell side >= cell_scale * rc_max */
    int32_t force_mode;    /* 0 = all pairs, 1 = cell list */
    int32_t n_threads;     /* OpenMP threads for the pair loops (<=0: runtime default) */
} orc_params;

typedef struct orc_obs {
    double K;              /* sum 1/2 m v^2 */
    double U;              /* shifted pair energy (+ wall energy in piston mode) */
    double W;              /* 1/2 sum_ordered r.F  (configurational virial) */
    double P;              /* instantaneous pressure as cal_instant_pressure() */
    double xi;             /* Nose-Hoover variable */
    double p_A;            /* NPH area momentum */
    double v_p;            /* piston velocity */
    double F_p;            /* piston force */
    double Lx, Ly;
    double t;
    double last_P;         /* get_last_instant_pressure() */
    double area_rate;      /* get_nph_area_rate() */
    double eta;            /* integral of xi dt over the NVT steps (Nose-Hoover conserved energy
                              K + U + Q xi^2 / 2 + g T0 eta; not in the reference) */
} orc_obs;

typedef struct orc_fire_params {
    double dt_init, dt_max, alpha0, f_inc, f_dec;
    int32_t N_min;
    int32_t energy_interval;   /* reference: 1000 */
    double max_step;           /* 0 = uncapped (reference behaviour) */
} orc_fire_params;

typedef struct orc_fire_result {
    int32_t steps;
    int32_t converged;
    double max_force;
    double potential_energy;
    int32_t n_energy;          /* number of (step, U) samples written */
} orc_fire_result;

typedef struct orc_sim orc_sim;

orc_sim *orc_create(const orc_params *p, const orc_particle *frame, int64_t n);
void orc_destroy(orc_sim *s);
void orc_get_frame(const orc_sim *s, orc_particle *out);
void orc_set_frame(orc_sim *s, const orc_particle *frame);
void orc_set_force_mode(orc_sim *s, int32_t mode);

/* Recompute accelerations at the current positions (periodic_y = 1 or 0). */
void orc_compute_forces(orc_sim *s, int32_t periodic_y);

/* Per-particle observables of one pair pass; any pointer may be NULL. */
void orc_pair_pass(const orc_sim *s, int32_t periodic_y, double *acc_xy,
                   double *u_i, double *w_i, double *lam_i, double eps_lambda);

double orc_total_K(const orc_sim *s);
double orc_total_U(orc_sim *s);
double orc_instant_pressure(orc_sim *s);
double orc_partial_U_lambda(orc_sim *s, double eps_lambda);
void orc_get_obs(const orc_sim *s, orc_obs *o);

void orc_step_nve(orc_sim *s, int32_t nsteps);
void orc_step_nvt(orc_sim *s, int32_t nsteps, int32_t middle_wrap);
void orc_step_nph(orc_sim *s, int32_t nsteps, double p_ext);
void orc_middle_reflect(orc_sim *s);
/* pressure-tensor profile in nb slabs along y: Pxx[nb], Pyy[nb], Pxy[nb] */
void orc_pressure_profile(orc_sim *s, int32_t nb, int32_t periodic_y, double *out);   /* MDSimulation::middle_reflect_LG alone */
void orc_step_piston(orc_sim *s, int32_t nsteps, double p_ext);

/* FIRE on the sim's current frame (wrapped, v = a = 0 first).  energy_steps /
 * energy_trace receive up to cap samples.  The frame is left in the sim. */
void orc_fire(orc_sim *s, const orc_fire_params *fp, double force_tol,
              int32_t max_steps, orc_fire_result *res, int32_t *energy_steps,
              double *energy_trace, int32_t cap);

/* Canonical cell list of the current positions.  keys[n], cell_start[ncell+1],
 * sorted[n].  Returns ncell; ncx/ncy out.  Arrays may be NULL to query sizes. */
int64_t orc_cell_list(const orc_sim *s, int32_t periodic_y, int32_t *ncx,
                      int32_t *ncy, uint32_t *keys, uint32_t *cell_start,
                      uint32_t *sorted);

/* Pure helpers exposed for known-answer tests. */
double orc_pbc_wrap(double x, double L);
void orc_pair(double dx, double dy, int32_t ti, int32_t tj, const orc_params *p,
              double *fx, double *fy, double *u, int32_t *in_cut);

#ifdef __cplusplus
}
#endif
