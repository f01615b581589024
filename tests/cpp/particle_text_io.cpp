// Writes the particle CSV and two XYZ frames of a 3-particle frame:
// particle_text_io <csv with box> <csv without box> <xyz>
#include <cstdio>
#include <vector>

#include "md_particle.hpp"

int main(int argc, char **argv) {
    if (argc != 4) return 2;
    std::vector<Particle> f(3);
    f[0].pos = make_double2(1.5, 2.25);
    f[0].type = 0;
    f[1].pos = make_double2(-0.125, 7.0);
    f[1].type = 1;
    f[2].pos = make_double2(3.0, 4.0);
    f[2].type = 5;
    print_particles_csv(f, argv[1], 10.0, 8.0, 1.0, 1.2);
    print_particles_csv(f, argv[2]);
    if (!append_xyz_frame(argv[3], f, "phase=NVT global_step=0 phase_step=0 global_time=0.0000000000 "
                                      "phase_time=0.0000000000 Lx=10.0000000000 Ly=8.0000000000"))
        return 3;
    if (!append_xyz_frame(argv[3], f, "second")) return 4;
    std::printf("ok\n");
    return 0;
}
