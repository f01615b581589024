// PFRAME v2 frame files (src/md_common.cpp:110-320 of the reference) and the
// small file helpers.  Frames are zlib compress2(level 9) of Particle[n] with
// a CRC32 of the uncompressed bytes; the 4 padding bytes of each record are
// written as zero so files are deterministic.
#include "md_common.hpp"

#include <cstring>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <vector>
#include <zlib.h>

bool create_folder(const std::filesystem::path &path, int rank_idx) {
    if (rank_idx != 0) return true;
    std::error_code ec;
    if (!std::filesystem::exists(path, ec)) {
        std::filesystem::create_directories(path, ec);
        if (ec) {
            std::cerr << "Failed to create dir " << path.string() << ". Error: " << ec.message() << "\n";
            return false;
        }
    }
    return true;
}

bool append_latest_line(const std::filesystem::path &src, const std::filesystem::path &dst, int rank_idx,
                        const std::string &tag) {
    if (rank_idx != 0) return true;
    std::ifstream in(src);
    if (!in) {
        std::cerr << "[" << tag << "] Failed to open " << src.string() << " for reading.\n";
        return false;
    }
    std::string line, last;
    while (std::getline(in, line))
        if (!line.empty()) last = line;
    if (last.empty()) {
        std::cerr << "[" << tag << "] No data found in " << src.string() << ".\n";
        return false;
    }
    std::ofstream out(dst, std::ios::out | std::ios::app);
    if (!out) {
        std::cerr << "[" << tag << "] Failed to open " << dst.string() << " for appending.\n";
        return false;
    }
    out << last << '\n';
    return true;
}

void write_exact(std::FILE *fp, const void *buf, std::size_t bytes) {
    if (bytes == 0) return;
    if (std::fwrite(buf, 1, bytes, fp) != bytes) throw std::runtime_error("Failed to write to file");
}

bool read_exact(std::FILE *fp, void *buf, std::size_t bytes) {
    if (bytes == 0) return true;
    const std::size_t got = std::fread(buf, 1, bytes, fp);
    return got == bytes;
}

FileWriter::FileWriter(const std::string &path, bool append) {
    fp_ = nullptr;
    if (append) {   // open existing, verify header, seek to end (src/md_common.cpp:112-135)
        fp_ = std::fopen(path.c_str(), "r+b");
        if (!fp_) throw std::runtime_error("Failed to open file for append: " + path);
        FileHeader header{};
        if (!read_exact(fp_, &header, sizeof(header))) {
            std::fclose(fp_);
            throw std::runtime_error("Failed to read existing header for append");
        }
        if (std::memcmp(header.magic, "PFRAME", 6) != 0 || (header.version != 1 && header.version != 2)) {
            std::fclose(fp_);
            throw std::runtime_error("Not a valid PFRAME v1/v2 file");
        }
        if (std::fseek(fp_, 0, SEEK_END) != 0) {
            std::fclose(fp_);
            throw std::runtime_error("Failed to seek to end of file");
        }
    } else {
        fp_ = std::fopen(path.c_str(), "wb");
        if (!fp_) throw std::runtime_error("Failed to open file for writing: " + path);
        FileHeader header{};
        std::memset(header.magic, 0, sizeof(header.magic));
        std::memcpy(header.magic, "PFRAME", 6);
        header.version = 2;
        header.reserved = 0;
        write_exact(fp_, &header, sizeof(header));
        std::fflush(fp_);
    }
}

FileWriter::~FileWriter() {
    if (fp_) {
        std::fflush(fp_);
        std::fclose(fp_);
    }
}

void FileWriter::write_frame(const Particle *data, std::size_t n_particles, std::uint64_t frame_index) {
    if (!fp_) throw std::runtime_error("File not open");
    const std::uint64_t uncompressed = static_cast<std::uint64_t>(n_particles) * sizeof(Particle);
    std::vector<unsigned char> raw(uncompressed);
    if (uncompressed) std::memcpy(raw.data(), data, uncompressed);
    for (std::size_t i = 0; i < n_particles; ++i)   // zero the padding after `type`
        std::memset(raw.data() + i * sizeof(Particle) + offsetof(Particle, type) + sizeof(int), 0, 4);
    uLong crc = ::crc32(0L, Z_NULL, 0);
    crc = ::crc32(crc, raw.data(), static_cast<uInt>(uncompressed));
    uLongf dest_len = ::compressBound(static_cast<uLong>(uncompressed));
    std::vector<unsigned char> comp(dest_len);
    if (::compress2(comp.data(), &dest_len, raw.data(), static_cast<uLong>(uncompressed), Z_BEST_COMPRESSION) != Z_OK)
        throw std::runtime_error("compress2 failed");
    FrameHeader fh{};
    fh.frame_index = frame_index;
    fh.n_particles = n_particles;
    fh.uncompressed_bytes = uncompressed;
    fh.compressed_bytes = dest_len;
    fh.crc32 = static_cast<std::uint32_t>(crc);
    write_exact(fp_, &fh, sizeof(fh));
    write_exact(fp_, comp.data(), dest_len);
    std::fflush(fp_);
}

FileReader::FileReader(const std::string &path) {
    fp_ = nullptr;
    eof_ = false;
    corrupted_ = false;
    version_ = 0;
    fp_ = std::fopen(path.c_str(), "rb");
    if (!fp_) throw std::runtime_error("Failed to open file for reading: " + path);
    FileHeader header{};
    if (!read_exact(fp_, &header, sizeof(header))) {
        std::fclose(fp_);
        throw std::runtime_error("Failed to read file header");
    }
    if (std::memcmp(header.magic, "PFRAME", 6) != 0) {
        std::fclose(fp_);
        throw std::runtime_error("Not a valid PFRAME file");
    }
    version_ = header.version;
    if (version_ == 1) {   // src/md_common.cpp:231-237
        std::fclose(fp_);
        throw std::runtime_error(
            "File format v1 is incompatible with current build (v2).\n"
            "  Reason: Particle struct now includes 'theta' field for ABP.\n"
            "  Solution: Regenerate simulation data with v2 format.");
    }
    if (version_ != 2) {
        std::fclose(fp_);
        throw std::runtime_error("Unknown file version: " + std::to_string(version_));
    }
}

FileReader::~FileReader() {
    if (fp_) std::fclose(fp_);
}

bool FileReader::eof() const { return eof_; }

bool FileReader::corrupted() const { return corrupted_; }

bool FileReader::next_frame(std::vector<Particle> &particles, std::uint64_t &frame_index_out) {
    if (!fp_ || eof_) return false;
    FrameHeader fh{};
    if (!read_exact(fp_, &fh, sizeof(fh))) {   // clean EOF or a broken partial header
        eof_ = true;
        return false;
    }
    if (fh.uncompressed_bytes != fh.n_particles * sizeof(Particle) || fh.compressed_bytes == 0) {
        corrupted_ = true;
        eof_ = true;
        return false;
    }
    std::vector<unsigned char> comp(static_cast<std::size_t>(fh.compressed_bytes));
    if (!read_exact(fp_, comp.data(), comp.size())) {
        corrupted_ = true;
        eof_ = true;
        return false;
    }
    std::vector<unsigned char> raw(static_cast<std::size_t>(fh.uncompressed_bytes));
    uLongf dest_len = static_cast<uLongf>(fh.uncompressed_bytes);
    const int zret = ::uncompress(raw.data(), &dest_len, comp.data(), static_cast<uLong>(fh.compressed_bytes));
    if (zret != Z_OK || dest_len != fh.uncompressed_bytes) {
        corrupted_ = true;
        eof_ = true;
        return false;
    }
    uLong crc = ::crc32(0L, Z_NULL, 0);
    crc = ::crc32(crc, raw.data(), static_cast<uInt>(fh.uncompressed_bytes));
    if (static_cast<std::uint32_t>(crc) != fh.crc32) {
        corrupted_ = true;
        eof_ = true;
        return false;
    }
    particles.resize(static_cast<std::size_t>(fh.n_particles));
    if (!raw.empty()) std::memcpy(particles.data(), raw.data(), raw.size());
    frame_index_out = fh.frame_index;
    return true;
}
