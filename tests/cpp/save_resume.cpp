// save_env and the resume constructor, as the reference's drivers use them
// (run/series_partial_U_lambda_test.cpp:283-288 saves with save_env after
// sample_collect; MDSimulation(cfg, comm, file, step), src/md_env.cu:81-140,
// resumes from the frame with that index).
//
//   save_resume run <dir>       (GPU) NVT from the constructor's frame in
//       windows of 16 list steps; frames saved at steps 40 and 60 into one
//       file; resumed simulations from both frames; checks below, then
//       middle_reflect_LG on a resumed simulation.
//   save_resume missing <dir>   resume from a frame index the file lacks
//   save_resume count <dir>     resume from a frame of the wrong size
//   save_resume badfile <dir>   resume from a file that is not PFRAME
// The last three fail before any device work: the process exits with status
// 1 and the reason on stderr, as the reference aborts.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <string>
#include <vector>

#include "md_env.hpp"

namespace {

MDConfigManager config(int n) {
    MDConfigManager m;
    MDConfig &c = m.config;
    c.n_particles_global = n;
    c.n_particles_type0 = n / 2;
    c.box_h_global = std::sqrt(n / 0.75 / 2.0);
    c.box_w_global = 2.0 * c.box_h_global;
    c.T_init = 0.6;
    c.T_target = 0.6;
    c.EPSILON_AB = 0.5;
    c.dt = 0.005;
    c.Q = 10.0;
    c.cutoff = 2.5;
    c.devide_p = 0.5;
    c.rank_size = 1;
    c.rank_idx = 0;
    return m;
}

std::vector<Particle> frame_of(MDSimulation &s) {
    std::vector<Particle> f;
    s.get_host_particles(f);
    return f;
}

// Fields equal bit for bit (the padding after `type` is not compared).
bool same_state(const std::vector<Particle> &a, const std::vector<Particle> &b, bool with_acc) {
    if (a.size() != b.size()) return false;
    for (std::size_t i = 0; i < a.size(); ++i) {
        const Particle &p = a[i], &q = b[i];
        if (std::memcmp(&p.pos, &q.pos, sizeof p.pos) || std::memcmp(&p.vel, &q.vel, sizeof p.vel) ||
            p.type != q.type || std::memcmp(&p.theta, &q.theta, sizeof p.theta))
            return false;
        if (with_acc && std::memcmp(&p.acc, &q.acc, sizeof p.acc)) return false;
    }
    return true;
}

double acc_rel_diff(const std::vector<Particle> &a, const std::vector<Particle> &b) {
    double ms = 0.0, mx = 0.0;
    for (std::size_t i = 0; i < a.size(); ++i) {
        ms += a[i].acc.x * a[i].acc.x + a[i].acc.y * a[i].acc.y;
        mx = std::fmax(mx, std::hypot(a[i].acc.x - b[i].acc.x, a[i].acc.y - b[i].acc.y));
    }
    return mx / std::sqrt(ms / static_cast<double>(a.size()));
}

int run_gpu(const std::filesystem::path &dir) {
    setenv("MDX_TUNING", "1,1,0.65,16", 1);   // fixed tuning: runs repeat bit for bit
    const int n = 4096;
    const std::string file = (dir / "env.bin").string();
    int bad = 0;
    auto expect = [&](bool ok, const char *what) {
        std::printf("%-62s %s\n", what, ok ? "ok" : "FAILED");
        bad += !ok;
    };
    std::vector<Particle> a40, a60;
    double U40 = 0.0, K40 = 0.0;
    {
        MDSimulation A(config(n), 0);
        for (int step = 0; step < 40; ++step) A.step_single_nose_hoover();
        A.sample_collect();
        A.save_env(file, 40);
        a40 = frame_of(A);
        U40 = A.cal_total_U();
        K40 = A.cal_total_K();
        for (int step = 40; step < 60; ++step) A.step_single_nose_hoover();
        A.sample_collect();
        A.save_env(file, 60);
        a60 = frame_of(A);
    }
    {
        FileReader r(file);
        std::vector<Particle> buf;
        std::uint64_t idx = 0;
        std::vector<std::uint64_t> seen;
        bool sizes = true;
        while (r.next_frame(buf, idx)) {
            seen.push_back(idx);
            sizes = sizes && buf.size() == static_cast<std::size_t>(n);
        }
        expect(seen == std::vector<std::uint64_t>{40, 60} && sizes && !r.corrupted(), "file holds frames 40 and 60");
    }
    MDSimulation B(config(n), 0, file, 40);
    expect(same_state(frame_of(B), a40, true), "resumed host frame == saved frame 40 (all fields)");
    B.sample_collect();
    const std::vector<Particle> b40 = frame_of(B);
    expect(same_state(b40, a40, false), "device round trip: pos, vel, type, theta bit-identical");
    const double dacc = acc_rel_diff(a40, b40);
    std::printf("  forces recomputed at load: max |da| / rms |a| = %.3e\n", dacc);
    expect(dacc <= 1e-5, "forces recomputed at load agree (FP32 summation order)");
    const double UB = B.cal_total_U(), KB = B.cal_total_K();
    std::printf("  U %.12e vs %.12e, K %.12e vs %.12e\n", UB, U40, KB, K40);
    expect(std::fabs(UB - U40) <= 1e-6 * std::fabs(U40) && std::fabs(KB - K40) <= 1e-12 * K40,
           "U and K of the resumed simulation");
    {
        MDSimulation C(config(n), 0, file, 60);
        expect(same_state(frame_of(C), a60, true), "resumed host frame == saved frame 60");
    }
    {   // a resume is reproducible: two resumes from one frame step alike
        MDSimulation B2(config(n), 0, file, 40);
        for (int step = 0; step < 30; ++step) {
            B.step_single_nose_hoover();
            B2.step_single_nose_hoover();
        }
        B.sample_collect();
        B2.sample_collect();
        expect(same_state(frame_of(B), frame_of(B2), true), "two resumes from frame 40: 30 NVT steps bit-identical");
    }
    {   // middle_reflect_LG through the class: every x ends in the central slab
        // of width devide_p Lx + 2 sigma_max (sigma = 1), every y in [0, Ly)
        B.middle_reflect_LG();
        B.sample_collect();
        const std::vector<Particle> f = frame_of(B);
        const double Lx = B.get_Lx(), Ly = B.get_Ly();
        const double slab = 0.5 * Lx + 2.0, x0 = 0.5 * Lx - 0.5 * slab;
        bool inside = true;
        int outside_before = 0;
        for (std::size_t i = 0; i < f.size(); ++i) {
            inside = inside && f[i].pos.x >= x0 && f[i].pos.x <= x0 + slab && f[i].pos.y >= 0.0 && f[i].pos.y < Ly;
            outside_before += a40[i].pos.x < x0 || a40[i].pos.x > x0 + slab;
        }
        std::printf("  middle_reflect_LG: %d particles were outside the slab\n", outside_before);
        expect(inside && outside_before > 0, "middle_reflect_LG folds every particle into the slab");
    }
    std::printf("save_resume bad=%d\n", bad);
    return bad == 0 ? 0 : 1;
}

// Failure paths: a small file, then a resume that must fail.
int run_failure(const std::string &mode, const std::filesystem::path &dir) {
    const std::string file = (dir / (mode + ".bin")).string();
    if (mode == "badfile") {
        std::ofstream(file, std::ios::binary) << "not a frame file at all";
    } else {
        std::vector<Particle> f(16);
        for (std::size_t i = 0; i < f.size(); ++i) f[i].pos = make_double2(1.0 + i, 2.0);
        FileWriter w(file, false);
        w.write_frame(f.data(), f.size(), 0);
        w.write_frame(f.data(), f.size(), 5);
    }
    MDConfigManager cfg = config(mode == "count" ? 20 : 16);
    MDSimulation s(cfg, 0, file, mode == "missing" ? 7 : 5);
    std::printf("resumed (unexpected)\n");
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: save_resume run|missing|count|badfile <dir>\n");
        return 2;
    }
    const std::string mode = argv[1];
    if (mode == "run") return run_gpu(argv[2]);
    return run_failure(mode, argv[2]);
}
