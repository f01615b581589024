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
