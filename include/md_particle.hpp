// Particle record of the reference (include/md_particle.hpp:31-38 there):
// 64 bytes, double2 pos @0, vel @16, acc @32 (acc = F/m), int type @48,
// 4 padding bytes, double theta @56.  It is also the PFRAME v2 payload.
#pragma once

#include <cstddef>

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
