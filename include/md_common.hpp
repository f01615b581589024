// PFRAME v2 frame files and small I/O helpers of the reference
// (include/md_common.hpp:17-122, src/md_common.cpp:110-320 there).
#pragma once

#include <cstdint>
#include <cstdio>
#include <filesystem>
#include <string>
#include <vector>

#include "md_particle.hpp"

#pragma pack(push, 1)
struct FileHeader {
    char magic[8];   // "PFRAME\0\0"
    std::uint32_t version;
    std::uint32_t reserved;
};

struct FrameHeader {
    std::uint64_t frame_index;
    std::uint64_t n_particles;
    std::uint64_t uncompressed_bytes;
    std::uint64_t compressed_bytes;
    std::uint32_t crc32;   // CRC of the uncompressed Particle data
};
#pragma pack(pop)

static_assert(sizeof(FileHeader) == 16, "FileHeader is 16 bytes");
static_assert(sizeof(FrameHeader) == 36, "FrameHeader is 36 bytes");

// Create a directory tree (rank 0 only); true on success.
bool create_folder(const std::filesystem::path &path, int rank_idx);

// Append the last non-empty line of src to dst (rank 0 only).
bool append_latest_line(const std::filesystem::path &src, const std::filesystem::path &dst,
                        int rank_idx, const std::string &tag);

// Rank 0 only: replace filepath with "bin,rho" and one "bin,value" line per
// entry, values in the default ostream format (6 significant digits); an
// open failure is reported on stderr under [tag].  Reference declaration:
// include/md_common.hpp:52-56.
void write_density_profile_csv(const std::filesystem::path &filepath, const std::vector<double> &density,
                               int rank_idx, const std::string &tag);

#if __has_include(<fmt/core.h>)
// Drivers built with {fmt}, as the reference's are, also get its two
// formatting helpers (include/md_common.hpp:33-40 and 63-82 there).
#include <fmt/core.h>
#include <fstream>
#include <utility>

template <typename... Args>
void RankZeroPrint(int rank_idx, fmt::format_string<Args...> format_str, Args &&...args) {
    if (rank_idx != 0) return;
    fmt::print(format_str, std::forward<Args>(args)...);
    std::fflush(stdout);
}

// Rank 0 only: append the formatted text (the caller ends it with '\n').
template <typename... Args>
bool append_csv(const std::filesystem::path &filepath, int rank_idx, const std::string &tag,
                fmt::format_string<Args...> fmt_str, Args &&...args) {
    if (rank_idx != 0) return true;
    const std::string text = fmt::format(fmt_str, std::forward<Args>(args)...);
    std::ofstream out(filepath, std::ios::out | std::ios::app);
    if (!out) {
        fmt::print(stderr, "[{}] Failed to open {} for appending.\n", tag, filepath.string());
        return false;
    }
    out << text;
    out.flush();
    return true;
}
#endif

// Write/read exactly this many bytes (write throws on a short write).
void write_exact(std::FILE *fp, const void *buf, std::size_t bytes);
bool read_exact(std::FILE *fp, void *buf, std::size_t bytes);

class FileWriter {
public:
    FileWriter(const std::string &path, bool append);
    ~FileWriter();
    void write_frame(const Particle *data, std::size_t n_particles, std::uint64_t frame_index);

private:
    std::FILE *fp_;
};

class FileReader {
public:
    explicit FileReader(const std::string &path);
    ~FileReader();
    bool eof() const;
    bool corrupted() const;
    bool next_frame(std::vector<Particle> &particles, std::uint64_t &frame_index_out);

private:
    std::FILE *fp_;
    bool eof_;
    bool corrupted_;
    std::uint32_t version_;
};
