// Device-resident FP64 scalars of one system and the kernel argument packs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "mdx_device.hpp"

namespace mdx {

// Global scalars of one system.  Written by single-thread box kernels or by
// the last workgroup of a grid reduction; read by later kernels.
struct DevScalars {
    double xi;        // Nose-Hoover variable after the last completed half update
    double xi_mid;    // xi after the first and second half updates of the step in flight
    double K;         // kinetic energy of the current velocities
    double K_half;    // kinetic energy after the first kick (NVT: v^{n+1/2}; NPH: scaled)
    double W;         // 1/2 sum_ordered r.F at the current positions
    double U;
    double lam;       // 1/2 sum_ordered (Fx dx - Fy dy) (unscaled)
    double wallE;
    double Lx, Ly;    // current box
    double p_A;       // NPH area momentum
    double gamma;     // NPH fixed aspect ratio
    double nph_lin, nph_vsc;
    double P_old;
    double v_p;       // piston velocity
    double F_p;       // piston force (upper wall)
    double last_P;
    double Ly_floor, Ly_ceil;
    int32_t clamped;
    int32_t box_error;
    // FIRE schedule (src/md_energy_minimizer.cu:170-203)
    double f_dt, f_alpha, f_maxF, f_power;
    int32_t f_streak, f_done, f_zero_v, f_step, f_conv_step, f_pad;
    unsigned ticket[4];
};

struct DriftArgs {
    double dt, T0, Q, g_dof;
    int middle_wrap;
    int pad;
    double slab_xmin, slab_period, slab_width;
    double max_step;
};

struct PairArgs {
    double dt;
    double p_ext;
    double W_baro;       // max(barostat_mass, 1e-12)
    double k_piston;
    // FIRE
    double ftol, f_inc, f_dec, dt_max, alpha0;
    int N_min;
    int pad;
};

}  // namespace mdx
