// Frame files and small file helpers with the reference's interface
// (include/md_common.hpp:44-122 there; behaviour of src/md_common.cpp:9-320).
//
// PFRAME v2 layout (little-endian, packed):
//   file   = FileHeader {char magic[8] = "PFRAME\0\0"; u32 version = 2; u32 0}
//            followed by frames until the end of the file;
//   frame  = FrameHeader {u64 index; u64 n; u64 raw = 64 n; u64 z; u32 crc}
//            followed by z bytes: the zlib stream (compress2, level 9) of the
//            n 64-byte Particle records, crc = CRC-32 of those raw bytes.
// The 4 padding bytes after `type` are written as zero, so equal frames give
// equal files.  A writer in append mode accepts a v1 or v2 file and adds
// frames at its end; a reader accepts v2 only.  A frame whose sizes disagree,
// whose bytes run short, or whose stream or CRC fails marks the reader
// corrupted and ends it.
#include "md_common.hpp"

#include <algorithm>
#include <cstddef>
#include <cstring>
#include <fstream>
#include <iostream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <vector>
#include <zlib.h>

namespace {

constexpr std::uint32_t kVersion = 2;
constexpr std::size_t kPadOffset = offsetof(Particle, type) + sizeof(int);   // bytes 52..55
constexpr std::size_t kPadBytes = offsetof(Particle, theta) - kPadOffset;
static_assert(sizeof(Particle) == 64 && kPadOffset == 52 && kPadBytes == 4, "64-byte record, padding 52..55");

using Bytes = std::vector<unsigned char>;

std::uint32_t crc_of(const Bytes &b) {
    uLong c = ::crc32(0L, Z_NULL, 0);
    const unsigned char *p = b.data();
    for (std::size_t left = b.size(); left > 0;) {   // crc32 takes uInt lengths
        const uInt part = static_cast<uInt>(std::min<std::size_t>(left, 1u << 30));
        c = ::crc32(c, p, part);
        p += part;
        left -= part;
    }
    return static_cast<std::uint32_t>(c);
}

// One frame as it goes to disk: FrameHeader then the compressed records.
Bytes encode_frame(const Particle *data, std::size_t n, std::uint64_t index) {
    Bytes raw(n * sizeof(Particle));
    if (n) std::memcpy(raw.data(), data, raw.size());
    for (std::size_t off = kPadOffset; off < raw.size(); off += sizeof(Particle))
        std::memset(raw.data() + off, 0, kPadBytes);
    uLongf zlen = ::compressBound(static_cast<uLong>(raw.size()));
    Bytes out(sizeof(FrameHeader) + zlen);
    if (::compress2(out.data() + sizeof(FrameHeader), &zlen, raw.data(), static_cast<uLong>(raw.size()),
                    Z_BEST_COMPRESSION) != Z_OK)
        throw std::runtime_error("compress2 failed");
    FrameHeader h{};
    h.frame_index = index;
    h.n_particles = n;
    h.uncompressed_bytes = raw.size();
    h.compressed_bytes = zlen;
    h.crc32 = crc_of(raw);
    std::memcpy(out.data(), &h, sizeof(h));
    out.resize(sizeof(FrameHeader) + zlen);
    return out;
}

bool has_magic(const FileHeader &h) { return std::memcmp(h.magic, "PFRAME", 6) == 0; }

struct FileCloser {
    void operator()(std::FILE *fp) const {
        if (fp) std::fclose(fp);
    }
};
using FileHandle = std::unique_ptr<std::FILE, FileCloser>;

}  // namespace

bool create_folder(const std::filesystem::path &path, int rank_idx) {
    if (rank_idx != 0) return true;
    std::error_code ec;
    if (std::filesystem::exists(path, ec)) return true;
    std::filesystem::create_directories(path, ec);
    if (!ec) return true;
    std::cerr << "Failed to create dir " << path.string() << ". Error: " << ec.message() << "\n";
    return false;
}

bool append_latest_line(const std::filesystem::path &src, const std::filesystem::path &dst, int rank_idx,
                        const std::string &tag) {
    if (rank_idx != 0) return true;
    std::ifstream in(src, std::ios::binary);
    if (!in) {
        std::cerr << "[" << tag << "] Failed to open " << src.string() << " for reading.\n";
        return false;
    }
    std::ostringstream whole;
    whole << in.rdbuf();
    const std::string text = whole.str();
    std::string last;   // the last '\n'-separated segment that is not empty
    for (std::size_t end = text.size();;) {
        const std::size_t nl = end == 0 ? std::string::npos : text.rfind('\n', end - 1);
        const std::size_t begin = nl == std::string::npos ? 0 : nl + 1;
        if (end > begin) {
            last = text.substr(begin, end - begin);
            break;
        }
        if (nl == std::string::npos) break;
        end = nl;
    }
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

void write_density_profile_csv(const std::filesystem::path &filepath, const std::vector<double> &density,
                               int rank_idx, const std::string &tag) {
    if (rank_idx != 0) return;
    std::ostringstream text;   // default stream formatting: 6 significant digits
    text << "bin,rho\n";
    for (std::size_t bin = 0; bin < density.size(); ++bin) text << bin << ',' << density[bin] << '\n';
    std::ofstream out(filepath, std::ios::out | std::ios::trunc);
    if (!out) {
        std::cerr << "[" << tag << "] Failed to open " << filepath.string() << " for writing density profile.\n";
        return;
    }
    out << text.str();
}

void write_exact(std::FILE *fp, const void *buf, std::size_t bytes) {
    const unsigned char *p = static_cast<const unsigned char *>(buf);
    while (bytes > 0) {
        const std::size_t done = std::fwrite(p, 1, bytes, fp);
        if (done == 0) throw std::runtime_error("Failed to write to file");
        p += done;
        bytes -= done;
    }
}

bool read_exact(std::FILE *fp, void *buf, std::size_t bytes) {
    unsigned char *p = static_cast<unsigned char *>(buf);
    while (bytes > 0) {
        const std::size_t got = std::fread(p, 1, bytes, fp);
        if (got == 0) return false;
        p += got;
        bytes -= got;
    }
    return true;
}

FileWriter::FileWriter(const std::string &path, bool append) : fp_(nullptr) {
    if (append) {
        FileHandle f(std::fopen(path.c_str(), "r+b"));
        if (!f) throw std::runtime_error("Failed to open file for append: " + path);
        FileHeader h{};
        if (!read_exact(f.get(), &h, sizeof(h))) throw std::runtime_error("Failed to read existing header for append");
        if (!has_magic(h) || (h.version != 1 && h.version != 2)) throw std::runtime_error("Not a valid PFRAME v1/v2 file");
        if (std::fseek(f.get(), 0, SEEK_END) != 0) throw std::runtime_error("Failed to seek to end of file");
        fp_ = f.release();
        return;
    }
    FileHandle f(std::fopen(path.c_str(), "wb"));
    if (!f) throw std::runtime_error("Failed to open file for writing: " + path);
    FileHeader h{};
    std::memcpy(h.magic, "PFRAME", 6);
    h.version = kVersion;
    write_exact(f.get(), &h, sizeof(h));
    std::fflush(f.get());
    fp_ = f.release();
}

FileWriter::~FileWriter() { FileHandle f(fp_); }   // fclose flushes

void FileWriter::write_frame(const Particle *data, std::size_t n_particles, std::uint64_t frame_index) {
    if (!fp_) throw std::runtime_error("File not open");
    const Bytes frame = encode_frame(data, n_particles, frame_index);
    write_exact(fp_, frame.data(), frame.size());
    std::fflush(fp_);
}

FileReader::FileReader(const std::string &path) : fp_(nullptr), eof_(false), corrupted_(false), version_(0) {
    FileHandle f(std::fopen(path.c_str(), "rb"));
    if (!f) throw std::runtime_error("Failed to open file for reading: " + path);
    FileHeader h{};
    if (!read_exact(f.get(), &h, sizeof(h))) throw std::runtime_error("Failed to read file header");
    if (!has_magic(h)) throw std::runtime_error("Not a valid PFRAME file");
    version_ = h.version;
    if (version_ == 1)   // the reference's message for v1 files (src/md_common.cpp:231-237)
        throw std::runtime_error(
            "File format v1 is incompatible with current build (v2).\n"
            "  Reason: Particle struct now includes 'theta' field for ABP.\n"
            "  Solution: Regenerate simulation data with v2 format.");
    if (version_ != kVersion) throw std::runtime_error("Unknown file version: " + std::to_string(version_));
    fp_ = f.release();
}

FileReader::~FileReader() { FileHandle f(fp_); }

bool FileReader::eof() const { return eof_; }

bool FileReader::corrupted() const { return corrupted_; }

bool FileReader::next_frame(std::vector<Particle> &particles, std::uint64_t &frame_index_out) {
    if (!fp_ || eof_) return false;
    FrameHeader h{};
    if (!read_exact(fp_, &h, sizeof(h))) {   // end of file (a partial header counts as the end)
        eof_ = true;
        return false;
    }
    // every failure below: the file is damaged from here on
    const bool sizes_ok = h.compressed_bytes > 0 && h.uncompressed_bytes == h.n_particles * sizeof(Particle);
    Bytes z, raw;
    bool ok = sizes_ok;
    if (ok) {
        z.resize(static_cast<std::size_t>(h.compressed_bytes));
        ok = read_exact(fp_, z.data(), z.size());
    }
    if (ok) {
        raw.resize(static_cast<std::size_t>(h.uncompressed_bytes));
        uLongf len = static_cast<uLongf>(raw.size());
        ok = ::uncompress(raw.data(), &len, z.data(), static_cast<uLong>(z.size())) == Z_OK && len == raw.size() &&
             crc_of(raw) == h.crc32;
    }
    if (!ok) {
        corrupted_ = true;
        eof_ = true;
        return false;
    }
    particles.resize(static_cast<std::size_t>(h.n_particles));
    if (!raw.empty()) std::memcpy(particles.data(), raw.data(), raw.size());
    frame_index_out = h.frame_index;
    return true;
}
