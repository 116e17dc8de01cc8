// Versioned on-disk alias table (no counterpart in the reference, which keeps
// tables in memory only: AliasTable.cuh:13-24, postBuild at
// AliasTableSplitGpu.cu:599-612).
//
// File layout, little-endian:
//   FileHeader (128 bytes): magic "ATGTABLE", format version, header size, n,
//     W, the build shape (layout, source dtype, PSA+, chains, section size,
//     split method, effective p), the caller's seed tag, the row checksum and
//     the header checksum;
//   n FP64 row weights;
//   n int32 aliases.
// Checksums: 64-bit FNV-1a style over little-endian 64-bit words (a trailing
// partial word is zero-padded), offset basis 14695981039346656037 and prime
// 1099511628211; rows: the weight bytes, then the alias bytes.
#include <cstdio>
#include <cstring>
#include <vector>

#include "atg_host.h"
#include "../../include/atg.h"

namespace atg {
namespace {

constexpr char kMagic[8] = {'A', 'T', 'G', 'T', 'A', 'B', 'L', 'E'};
constexpr uint32_t kVersion = 1;

struct FileHeader {
    char magic[8];
    uint32_t version;
    uint32_t header_bytes;
    int64_t n;
    double W;
    int32_t layout;
    int32_t source_dtype;
    int32_t psa_plus;
    int32_t chains;
    int32_t split_method;
    int32_t reserved0;
    int64_t section_size;
    int64_t num_splits;
    uint64_t seed;
    uint64_t rows_checksum;
    uint64_t header_checksum;  // over every byte before this field
    uint8_t pad[32];
};
static_assert(sizeof(FileHeader) == 128, "file header is 128 bytes");

struct Fnv {
    uint64_t h = 14695981039346656037ull;
    void words(const void* p, size_t bytes) {
        const unsigned char* b = static_cast<const unsigned char*>(p);
        size_t i = 0;
        for (; i + 8 <= bytes; i += 8) {
            uint64_t w;
            std::memcpy(&w, b + i, 8);
            h = (h ^ w) * 1099511628211ull;
        }
        if (i < bytes) {
            uint64_t w = 0;
            std::memcpy(&w, b + i, bytes - i);
            h = (h ^ w) * 1099511628211ull;
        }
    }
};

uint64_t rows_checksum(const double* w, const int32_t* al, int64_t n) {
    Fnv f;
    f.words(w, sizeof(double) * (size_t)n);
    f.words(al, sizeof(int32_t) * (size_t)n);
    return f.h;
}

uint64_t header_checksum(const FileHeader& h) {
    Fnv f;
    f.words(&h, offsetof(FileHeader, header_checksum));
    return f.h;
}

struct File {
    FILE* f = nullptr;
    ~File() {
        if (f) std::fclose(f);
    }
};

}  // namespace
}  // namespace atg

using namespace atg;

extern "C" {

atg_status atg_table_file_write(const char* path, const atg_table_file_info* info, const double* weights,
                                const int32_t* aliases) {
    if (!path || !info || !weights || !aliases) return fail(ATG_EINVAL, "atg_table_file_write: NULL argument");
    if (info->n < 1) return fail(ATG_EINVAL, "atg_table_file_write: n=%lld", (long long)info->n);
    FileHeader h;
    std::memset(&h, 0, sizeof(h));
    std::memcpy(h.magic, kMagic, 8);
    h.version = kVersion;
    h.header_bytes = sizeof(FileHeader);
    h.n = info->n;
    h.W = info->W;
    h.layout = info->layout;
    h.source_dtype = info->source_dtype;
    h.psa_plus = info->psa_plus;
    h.chains = info->chains;
    h.split_method = info->split_method;
    h.section_size = info->section_size;
    h.num_splits = info->num_splits;
    h.seed = info->seed;
    h.rows_checksum = rows_checksum(weights, aliases, info->n);
    h.header_checksum = header_checksum(h);
    File out;
    out.f = std::fopen(path, "wb");
    if (!out.f) return fail(ATG_EIO, "atg_table_file_write: cannot open %s", path);
    const size_t n = (size_t)info->n;
    if (std::fwrite(&h, sizeof(h), 1, out.f) != 1 || std::fwrite(weights, sizeof(double), n, out.f) != n ||
        std::fwrite(aliases, sizeof(int32_t), n, out.f) != n)
        return fail(ATG_EIO, "atg_table_file_write: write to %s failed", path);
    if (std::fclose(out.f) != 0) {
        out.f = nullptr;
        return fail(ATG_EIO, "atg_table_file_write: closing %s failed", path);
    }
    out.f = nullptr;
    return ATG_OK;
}

atg_status atg_table_file_read(const char* path, atg_table_file_info* info, double* weights, int32_t* aliases,
                               int64_t capacity) {
    if (!path || !info) return fail(ATG_EINVAL, "atg_table_file_read: NULL argument");
    File in;
    in.f = std::fopen(path, "rb");
    if (!in.f) return fail(ATG_EIO, "atg_table_file_read: cannot open %s", path);
    FileHeader h;
    if (std::fread(&h, sizeof(h), 1, in.f) != 1) return fail(ATG_EIO, "atg_table_file_read: %s: short header", path);
    if (std::memcmp(h.magic, kMagic, 8) != 0) return fail(ATG_EIO, "atg_table_file_read: %s is not a table file", path);
    if (h.version != kVersion || h.header_bytes != sizeof(FileHeader))
        return fail(ATG_ENOTSUP, "atg_table_file_read: %s: format version %u (header %u bytes) not supported", path,
                    h.version, h.header_bytes);
    if (h.header_checksum != header_checksum(h)) return fail(ATG_EIO, "atg_table_file_read: %s: header checksum mismatch", path);
    if (h.n < 1) return fail(ATG_EIO, "atg_table_file_read: %s: n=%lld", path, (long long)h.n);
    info->n = h.n;
    info->W = h.W;
    info->layout = h.layout;
    info->source_dtype = h.source_dtype;
    info->psa_plus = h.psa_plus;
    info->chains = h.chains;
    info->split_method = h.split_method;
    info->reserved = 0;
    info->section_size = h.section_size;
    info->num_splits = h.num_splits;
    info->seed = h.seed;
    info->checksum = h.rows_checksum;
    if (!weights && !aliases) return ATG_OK;  // header only
    if (!weights || !aliases) return fail(ATG_EINVAL, "atg_table_file_read: weights and aliases go together");
    if (capacity < h.n)
        return fail(ATG_EINVAL, "atg_table_file_read: capacity %lld < n=%lld", (long long)capacity, (long long)h.n);
    const size_t n = (size_t)h.n;
    if (std::fread(weights, sizeof(double), n, in.f) != n || std::fread(aliases, sizeof(int32_t), n, in.f) != n)
        return fail(ATG_EIO, "atg_table_file_read: %s: truncated rows", path);
    if (rows_checksum(weights, aliases, h.n) != h.rows_checksum)
        return fail(ATG_EIO, "atg_table_file_read: %s: row checksum mismatch", path);
    return ATG_OK;
}

}  // extern "C"
