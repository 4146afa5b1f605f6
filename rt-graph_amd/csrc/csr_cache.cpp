// On-disk cache of a preprocessed CSR (checkpoint / resume of the ingest and preprocessing
// steps). The reference's closest mechanism is its text -> binary edge cache
// (tc/graph/graph.cpp:179-214); here the oriented (or symmetric) CSR itself is stored, so a
// later run skips the relabel, orientation, sort and dedup.
//
// Layout, little-endian:
//   char     magic[8]     "RTGCSR1\0"
//   uint32   version      1
//   uint32   flags        bit 0: degree-oriented
//   int32    n
//   int32    reserved     0
//   int64    m
//   uint64   key          caller-defined identity of the input (e.g. rtg_pairs_fingerprint of
//                         the pair list and the preprocessing flags); a load with another key
//                         is refused as stale
//   uint64   checksum     word-wise FNV-1a over offsets[n+1] then adjs[m]
//   int32    offsets[n+1], int32 adjs[m]
#include "rtg_internal.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

constexpr char kMagic[8] = {'R', 'T', 'G', 'C', 'S', 'R', '1', '\0'};
constexpr uint32_t kVersion = 1;

struct Header {
    char magic[8];
    uint32_t version;
    uint32_t flags;
    int32_t n;
    int32_t reserved;
    int64_t m;
    uint64_t key;
    uint64_t checksum;
};
static_assert(sizeof(Header) == 48, "packed header");

// FNV-1a over 64-bit little-endian words (a trailing partial word zero-padded), continuing
// from h. Byte-wise FNV would cost several times more on arrays of hundreds of MB.
uint64_t fnv1a_words(uint64_t h, const void *data, size_t bytes) {
    const unsigned char *p = static_cast<const unsigned char *>(data);
    size_t i = 0;
    for (; i + 8 <= bytes; i += 8) {
        uint64_t w;
        std::memcpy(&w, p + i, 8);
        h = (h ^ w) * 0x100000001B3ull;
    }
    if (i < bytes) {
        uint64_t w = 0;
        std::memcpy(&w, p + i, bytes - i);
        h = (h ^ w) * 0x100000001B3ull;
    }
    return h;
}

constexpr uint64_t kFnvBasis = 0xCBF29CE484222325ull;

uint64_t csr_checksum(const int32_t *offsets, const int32_t *adjs, int32_t n, int64_t m) {
    uint64_t h = fnv1a_words(kFnvBasis, offsets, sizeof(int32_t) * ((size_t)n + 1));
    if (m > 0) h = fnv1a_words(h, adjs, sizeof(int32_t) * (size_t)m);
    return h;
}

struct File {
    FILE *f = nullptr;
    ~File() {
        if (f) std::fclose(f);
    }
};

}  // namespace

extern "C" {

RTG_API int rtg_pairs_fingerprint(const int32_t *pairs, int64_t npairs, int32_t flags, uint64_t *out) {
    if (!out || npairs < 0 || (npairs > 0 && !pairs)) return rtg_fail(RTG_EINVAL, "rtg_pairs_fingerprint: bad arguments");
    uint64_t h = fnv1a_words(kFnvBasis, &flags, sizeof(flags));
    h = fnv1a_words(h, &npairs, sizeof(npairs));
    if (npairs > 0) h = fnv1a_words(h, pairs, sizeof(int32_t) * 2 * (size_t)npairs);
    *out = h;
    return RTG_OK;
}

RTG_API int rtg_csr_file_write(const char *path, uint64_t key, int32_t oriented, const int32_t *offsets,
                               const int32_t *adjs, int32_t n, int64_t m) {
    if (!path || !offsets || n < 0 || m < 0 || (m > 0 && !adjs))
        return rtg_fail(RTG_EINVAL, "rtg_csr_file_write: bad arguments");
    if (offsets[0] != 0 || offsets[n] != m) return rtg_fail(RTG_EINVAL, "rtg_csr_file_write: offsets[0] != 0 or offsets[n] != m");
    Header h{};
    std::memcpy(h.magic, kMagic, sizeof(kMagic));
    h.version = kVersion;
    h.flags = oriented ? 1u : 0u;
    h.n = n;
    h.m = m;
    h.key = key;
    h.checksum = csr_checksum(offsets, adjs, n, m);
    // Written to a temporary name and renamed, so that a reader never sees a partial file.
    const std::string tmp = std::string(path) + ".part";
    {
        File f;
        f.f = std::fopen(tmp.c_str(), "wb");
        if (!f.f) return rtg_fail(RTG_EIO, std::string("rtg_csr_file_write: cannot create ") + tmp);
        bool ok = std::fwrite(&h, sizeof(h), 1, f.f) == 1 &&
                  std::fwrite(offsets, sizeof(int32_t), (size_t)n + 1, f.f) == (size_t)n + 1 &&
                  (m == 0 || std::fwrite(adjs, sizeof(int32_t), (size_t)m, f.f) == (size_t)m);
        ok = (std::fclose(f.f) == 0) && ok;
        f.f = nullptr;
        if (!ok) {
            std::remove(tmp.c_str());
            return rtg_fail(RTG_EIO, std::string("rtg_csr_file_write: write failed: ") + tmp);
        }
    }
    if (std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        return rtg_fail(RTG_EIO, std::string("rtg_csr_file_write: cannot rename to ") + path);
    }
    return RTG_OK;
}

RTG_API int rtg_csr_file_read(const char *path, uint64_t key, int32_t *offsets, int32_t *adjs, int32_t *n,
                              int64_t *m, int32_t *oriented) {
    if (!path || !n || !m) return rtg_fail(RTG_EINVAL, "rtg_csr_file_read: bad arguments");
    File f;
    f.f = std::fopen(path, "rb");
    if (!f.f) return rtg_fail(RTG_EIO, std::string("rtg_csr_file_read: cannot open ") + path);
    Header h{};
    if (std::fread(&h, sizeof(h), 1, f.f) != 1 || std::memcmp(h.magic, kMagic, sizeof(kMagic)) != 0)
        return rtg_fail(RTG_EIO, "rtg_csr_file_read: not a CSR cache file");
    if (h.version != kVersion) return rtg_fail(RTG_EIO, "rtg_csr_file_read: unsupported version");
    if (h.key != key) return rtg_fail(RTG_EIO, "rtg_csr_file_read: the file was made from another input (key differs)");
    if (h.n < 0 || h.m < 0 || h.m >= ((int64_t)1 << 31)) return rtg_fail(RTG_EIO, "rtg_csr_file_read: bad sizes");
    // the file must hold exactly the header and both arrays
    if (std::fseek(f.f, 0, SEEK_END) != 0) return rtg_fail(RTG_EIO, "rtg_csr_file_read: cannot seek");
    const long size = std::ftell(f.f);
    const int64_t want = (int64_t)sizeof(h) + 4 * ((int64_t)h.n + 1) + 4 * h.m;
    if (size != want) return rtg_fail(RTG_EIO, "rtg_csr_file_read: truncated or oversized file");
    if (!offsets) {  // first call of the two-call convention: sizes only
        *n = h.n;
        *m = h.m;
        if (oriented) *oriented = (int32_t)(h.flags & 1u);
        return RTG_OK;
    }
    if (*n < h.n || *m < h.m || (h.m > 0 && !adjs)) return rtg_fail(RTG_EINVAL, "rtg_csr_file_read: buffers too small");
    if (std::fseek(f.f, (long)sizeof(h), SEEK_SET) != 0 ||
        std::fread(offsets, sizeof(int32_t), (size_t)h.n + 1, f.f) != (size_t)h.n + 1 ||
        (h.m > 0 && std::fread(adjs, sizeof(int32_t), (size_t)h.m, f.f) != (size_t)h.m))
        return rtg_fail(RTG_EIO, "rtg_csr_file_read: short read");
    if (csr_checksum(offsets, adjs, h.n, h.m) != h.checksum) return rtg_fail(RTG_EIO, "rtg_csr_file_read: checksum mismatch");
    *n = h.n;
    *m = h.m;
    if (oriented) *oriented = (int32_t)(h.flags & 1u);
    return RTG_OK;
}

RTG_API int rtg_graph_save(const rtg_graph *g, const char *path, uint64_t key) {
    if (!g || !path) return rtg_fail(RTG_EINVAL, "rtg_graph_save: null argument");
    int32_t n = 0, max_row = 0;
    int64_t m = 0;
    if (const int rc = rtg_graph_info(g, &n, &m, &max_row)) return rc;
    std::vector<int32_t> off((size_t)n + 1), adj((size_t)std::max<int64_t>(m, 1));
    if (const int rc = rtg_graph_download(g, off.data(), adj.data())) return rc;
    return rtg_csr_file_write(path, key, rtg::graph_oriented(g) ? 1 : 0, off.data(), adj.data(), n, m);
}

RTG_API int rtg_graph_load(rtg_ctx *ctx, const char *path, uint64_t key, rtg_graph **out) {
    if (!ctx || !path || !out) return rtg_fail(RTG_EINVAL, "rtg_graph_load: null argument");
    int32_t n = 0, oriented = 1;
    int64_t m = 0;
    if (const int rc = rtg_csr_file_read(path, key, nullptr, nullptr, &n, &m, &oriented)) return rc;
    std::vector<int32_t> off((size_t)n + 1), adj((size_t)std::max<int64_t>(m, 1));
    if (const int rc = rtg_csr_file_read(path, key, off.data(), adj.data(), &n, &m, &oriented)) return rc;
    if (const int rc = rtg_graph_from_csr(ctx, off.data(), adj.data(), n, m, out)) return rc;
    rtg::graph_set_oriented(*out, oriented != 0);
    return RTG_OK;
}

}  // extern "C"
