// PFRAME round trip with the shapes of the reference's
// tests/save_load_frame/save_load_frame.cpp: a 3-particle frame 0 and a
// 2-particle frame 1, read back and compared; then append mode and the
// corrupted-tail and v1-rejection paths.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "md_common.hpp"

static bool same(const Particle &a, const Particle &b) {
    return a.pos.x == b.pos.x && a.pos.y == b.pos.y && a.vel.x == b.vel.x && a.vel.y == b.vel.y &&
           a.acc.x == b.acc.x && a.acc.y == b.acc.y && a.type == b.type && a.theta == b.theta;
}

int main(int argc, char **argv) {
    if (argc < 2) return 2;
    const std::string path = argv[1];
    std::vector<Particle> f0(3), f1(2);
    for (std::size_t i = 0; i < f0.size(); i++) {
        std::memset(&f0[i], 0xAB, sizeof(Particle));   // padding must not leak
        f0[i].pos = {double(i), double(i) * 2.0};
        f0[i].vel = {0.1, 0.2};
        f0[i].acc = {0.0, 0.0};
        f0[i].type = int(i);
        f0[i].theta = 0.5 * i;
    }
    for (std::size_t i = 0; i < f1.size(); i++) {
        std::memset(&f1[i], 0, sizeof(Particle));
        f1[i].pos = {double(i), double(i) * 3.0};
        f1[i].vel = {0.3, 0.4};
        f1[i].type = int(10 + i);
    }
    int bad = 0;
    {
        FileWriter w(path, false);
        w.write_frame(f0.data(), f0.size(), 0);
        w.write_frame(f1.data(), f1.size(), 1);
    }
    {
        std::vector<Particle> f2(1);
        std::memset(&f2[0], 0, sizeof(Particle));
        f2[0].pos = {7.0, 8.0};
        FileWriter w(path, true);   // append mode verifies the header
        w.write_frame(f2.data(), f2.size(), 5);
    }
    {
        FileReader r(path);
        std::vector<Particle> buf;
        std::uint64_t idx = 0;
        int nf = 0;
        while (r.next_frame(buf, idx)) {
            const std::vector<Particle> &want = nf == 0 ? f0 : f1;
            if (nf < 2) {
                bad += buf.size() != want.size();
                for (std::size_t i = 0; i < buf.size() && i < want.size(); ++i) bad += !same(buf[i], want[i]);
                bad += idx != static_cast<std::uint64_t>(nf);
            } else {
                bad += idx != 5 || buf.size() != 1 || buf[0].pos.x != 7.0;
            }
            ++nf;
        }
        bad += nf != 3;
        bad += r.corrupted();
        bad += !r.eof();
    }
    // truncated tail: drop the last 5 bytes
    {
        std::FILE *fp = std::fopen(path.c_str(), "rb");
        std::vector<unsigned char> all;
        int c;
        while ((c = std::fgetc(fp)) != EOF) all.push_back(static_cast<unsigned char>(c));
        std::fclose(fp);
        const std::string tpath = path + ".trunc";
        fp = std::fopen(tpath.c_str(), "wb");
        std::fwrite(all.data(), 1, all.size() - 5, fp);
        std::fclose(fp);
        FileReader r(tpath);
        std::vector<Particle> buf;
        std::uint64_t idx = 0;
        int nf = 0;
        while (r.next_frame(buf, idx)) ++nf;
        bad += nf != 2 || !r.corrupted();
        // version 1 header is rejected
        const std::string vpath = path + ".v1";
        fp = std::fopen(vpath.c_str(), "wb");
        FileHeader h{};
        std::memcpy(h.magic, "PFRAME", 6);
        h.version = 1;
        std::fwrite(&h, sizeof(h), 1, fp);
        std::fclose(fp);
        bool threw = false;
        try {
            FileReader r1(vpath);
        } catch (const std::exception &e) {
            threw = std::string(e.what()).find("v1 is incompatible") != std::string::npos;
        }
        bad += !threw;
    }
    std::printf("save_load_frame bad=%d\n", bad);
    return bad == 0 ? 0 : 1;
}
