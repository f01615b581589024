// The text helpers of md_common.hpp as the reference's drivers use them:
// common_text_io <dir>
//   <dir>/made/a/b              create_folder on rank 0
//   <dir>/not_made              create_folder on rank 1 (nothing created)
//   <dir>/density.csv           write_density_profile_csv over a file that
//                               already holds text (it is replaced)
//   <dir>/density_rank1.csv     rank 1: not written
//   <dir>/latest.csv            append_latest_line twice from src.csv
// Exit code 0 when every return value is the expected one.
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "md_common.hpp"

int main(int argc, char **argv) {
    if (argc != 2) return 2;
    const std::filesystem::path dir = argv[1];
    int bad = 0;
    bad += !create_folder(dir / "made" / "a" / "b", 0);
    bad += !create_folder(dir / "made" / "a" / "b", 0);   // already there
    bad += !create_folder(dir / "not_made", 1);
    {
        std::ofstream old(dir / "density.csv");
        old << "stale contents that must disappear\n1,2,3\n4,5,6\n7,8,9\n";
    }
    const std::vector<double> rho = {0.5, -1.0 / 3.0, 1e-7, 123456789.0, 0.0, -0.25, 1.0, 2.0 / 3.0, 1e300};
    write_density_profile_csv(dir / "density.csv", rho, 0, "common_text_io");
    write_density_profile_csv(dir / "density_rank1.csv", rho, 1, "common_text_io");
    write_density_profile_csv(dir / "empty.csv", {}, 0, "common_text_io");
    {
        std::ofstream src(dir / "src.csv");
        src << "a,1\nb,2\n\n";
    }
    {
        std::ofstream src(dir / "src_noeol.csv");
        src << "c,3\nd,4";
    }
    {
        std::ofstream src(dir / "blank.csv");
        src << "\n\n";
    }
    bad += !append_latest_line(dir / "src.csv", dir / "latest.csv", 0, "common_text_io");
    bad += !append_latest_line(dir / "src_noeol.csv", dir / "latest.csv", 0, "common_text_io");
    bad += !append_latest_line(dir / "src.csv", dir / "latest_rank1.csv", 1, "common_text_io");
    bad += append_latest_line(dir / "blank.csv", dir / "latest.csv", 0, "common_text_io");     // no data: false
    bad += append_latest_line(dir / "missing.csv", dir / "latest.csv", 0, "common_text_io");   // no file: false
    std::printf("common_text_io bad=%d\n", bad);
    return bad == 0 ? 0 : 1;
}
