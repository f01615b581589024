// Particle record of the reference (include/md_particle.hpp:31-38 there):
// 64 bytes, double2 pos @0, vel @16, acc @32 (acc = F/m), int type @48,
// 4 padding bytes, double theta @56.  It is also the PFRAME v2 payload.
#pragma once

#include <cstddef>
#include <string>
#include <vector>

#include "mdx.h"

// Host stand-in for the CUDA vector type the reference's drivers use.
struct alignas(16) double2 {
    double x;
    double y;
};

inline double2 make_double2(double x, double y) { return double2{x, y}; }

struct Particle {
    double2 pos;
    double2 vel;
    double2 acc;
    int type;
    double theta;   // ABP orientation; carried through, not integrated here
};

static_assert(sizeof(Particle) == 64, "Particle must stay 64 bytes");
static_assert(offsetof(Particle, vel) == 16 && offsetof(Particle, acc) == 32 &&
                  offsetof(Particle, type) == 48 && offsetof(Particle, theta) == 56,
              "Particle layout must match the PFRAME v2 payload");
static_assert(sizeof(Particle) == sizeof(mdx_particle64), "Particle and mdx_particle64 differ");

// Particle CSV snapshot (reference src/md_particle.cpp:34-86,
// include/md_particle.hpp:40-46): a metadata row "box_w,<w>,box_h,<h>,
// sigma_aa,<s>,sigma_bb,<s>,draw_box,<0|1>,n_particles,<n>", then one row
// "x,<x>,y,<y>,type,<t>" per particle, fixed notation with 12 digits.  A box
// side <= 0 is replaced by the particles' extent along it (1 when that is
// not positive); draw_box is 1 only when both sides were given.  Throws
// std::runtime_error when the file cannot be opened or written.
void print_particles_csv(const std::vector<Particle> &particles, const std::string &filename,
                         const double box_w = 0.0, const double box_h = 0.0, const double sigma_aa = 1.0,
                         const double sigma_bb = 1.0);

// One XYZ frame appended to `path` (the snapshot format of the reference's
// NPH batch driver, run/series_NPH_batch_xyz_saving.cpp:106-141): the
// particle count, the comment line, then "<A|B|X> <x> <y> 0.0" per particle
// (type 0 -> A, 1 -> B, any other -> X), fixed notation with 10 digits.
// The driver's comment is "phase=<NVT|NPH> global_step=<i> phase_step=<i>
// global_time=<t> phase_time=<t> Lx=<L> Ly=<L>".  Returns false when the
// file cannot be opened or written.
bool append_xyz_frame(const std::string &path, const std::vector<Particle> &particles,
                      const std::string &comment);
