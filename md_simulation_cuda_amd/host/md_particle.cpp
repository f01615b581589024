// Text snapshots of a frame: the reference's particle CSV
// (src/md_particle.cpp:34-86) and the XYZ frames of its NPH batch driver
// (run/series_NPH_batch_xyz_saving.cpp:106-141).
#include "md_particle.hpp"

#include <algorithm>
#include <fstream>
#include <iomanip>
#include <stdexcept>

namespace {

// side of the box to write: the given one, else the particles' extent along
// that axis, else 1
double resolved_side(double given, const std::vector<Particle> &particles, bool along_x) {
    if (given > 0.0) return given;
    if (particles.empty()) return 1.0;
    double lo = along_x ? particles.front().pos.x : particles.front().pos.y;
    double hi = lo;
    for (const Particle &p : particles) {
        const double c = along_x ? p.pos.x : p.pos.y;
        lo = std::min(lo, c);
        hi = std::max(hi, c);
    }
    return hi - lo > 0.0 ? hi - lo : 1.0;
}

const char *xyz_label(int type) { return type == 0 ? "A" : (type == 1 ? "B" : "X"); }

}  // namespace

void print_particles_csv(const std::vector<Particle> &particles, const std::string &filename, const double box_w,
                         const double box_h, const double sigma_aa, const double sigma_bb) {
    const bool draw_box = box_w > 0.0 && box_h > 0.0;
    const double w = resolved_side(box_w, particles, true);
    const double h = resolved_side(box_h, particles, false);
    std::ofstream out(filename, std::ios::trunc);
    if (!out) throw std::runtime_error("Failed to open particle CSV for writing: " + filename);
    out << std::fixed << std::setprecision(12);
    out << "box_w," << w << ",box_h," << h << ",sigma_aa," << sigma_aa << ",sigma_bb," << sigma_bb
        << ",draw_box," << (draw_box ? 1 : 0) << ",n_particles," << particles.size() << '\n';
    for (const Particle &p : particles) out << "x," << p.pos.x << ",y," << p.pos.y << ",type," << p.type << '\n';
    if (!out) throw std::runtime_error("Failed to write particle CSV: " + filename);
}

bool append_xyz_frame(const std::string &path, const std::vector<Particle> &particles, const std::string &comment) {
    std::ofstream out(path, std::ios::out | std::ios::app);
    if (!out) return false;
    out << std::fixed << std::setprecision(10);
    out << particles.size() << '\n' << comment << '\n';
    for (const Particle &p : particles) out << xyz_label(p.type) << ' ' << p.pos.x << ' ' << p.pos.y << " 0.0\n";
    out.flush();
    return static_cast<bool>(out);
}
