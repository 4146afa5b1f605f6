// Host-side edge-list ingest in the reference's three on-disk formats.
//
//   binary  little-endian int32 (src, dst) pairs, npairs = file_size / 8
//           (Graph::ReadBinaryEdges, tc/graph/graph.cpp:23-39)
//   text    one "src dst" pair per line, unparsable lines skipped with a message
//           (Graph::Txt2Bin, tc/graph/graph.cpp:179-214)
//   .mtx    Matrix Market coordinate pattern/real/integer, 1-based, converted to 0-based;
//           "symmetric" files get every off-diagonal entry twice, (r,c) then (c,r)
//           (Graph::ReadMtx, bfs/graph/graph.cpp:17-122)
//
// All readers use the two-call convention: pairs == NULL returns the pair count only.
#include "rtg_internal.hpp"

#include <cctype>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace {

int read_text_pairs(const char *path, std::vector<int32_t> &out) {
    std::ifstream in(path);
    if (!in) return rtg_fail(RTG_EIO, std::string("cannot open ") + path);
    std::string line;
    while (std::getline(in, line)) {
        std::stringstream ss(line);
        int32_t a, b;  // as the reference's `int` extraction: an out-of-range value fails the line
        if (ss >> a >> b) {
            out.push_back(a);
            out.push_back(b);
        } else if (!line.empty() && line[0] != '#' && line[0] != '%') {
            std::fprintf(stderr, "Error parsing line: %s\n", line.c_str());
        }
    }
    return RTG_OK;
}

int read_mtx_pairs(const char *path, std::vector<int32_t> &out, int32_t *rows_out) {
    FILE *f = std::fopen(path, "r");
    if (!f) return rtg_fail(RTG_EIO, std::string("File could not be opened: ") + path);
    struct Closer {
        FILE *f;
        ~Closer() { std::fclose(f); }
    } closer{f};
    char buf[1024];
    if (!std::fgets(buf, sizeof(buf), f)) return rtg_fail(RTG_EIO, "Could not process Matrix Market banner");
    std::string banner(buf);
    for (auto &c : banner) c = (char)std::tolower((unsigned char)c);
    if (banner.rfind("%%matrixmarket", 0) != 0) return rtg_fail(RTG_EIO, "Could not process Matrix Market banner");
    if (banner.find("matrix") == std::string::npos || banner.find("array") != std::string::npos)
        return rtg_fail(RTG_EIO, "File is not a sparse matrix");
    const bool pattern = banner.find("pattern") != std::string::npos;
    const bool real = banner.find("real") != std::string::npos || banner.find("integer") != std::string::npos;
    const bool symmetric = banner.find("symmetric") != std::string::npos;
    if (!pattern && !real) return rtg_fail(RTG_EIO, "Unrecognized matrix market format type");
    long long rows = 0, cols = 0, nnz = 0;
    bool have_size = false;
    while (std::fgets(buf, sizeof(buf), f)) {
        if (buf[0] == '%') continue;
        if (std::sscanf(buf, "%lld %lld %lld", &rows, &cols, &nnz) == 3) {
            have_size = true;
            break;
        }
        return rtg_fail(RTG_EIO, "Could not read file info (M, N, NNZ)");
    }
    if (!have_size || rows < 0 || cols < 0 || nnz < 0 || rows > INT32_MAX || cols > INT32_MAX)
        return rtg_fail(RTG_EIO, "Could not read file info (M, N, NNZ)");
    out.reserve((size_t)nnz * (symmetric ? 4 : 2));
    for (long long i = 0; i < nnz; ++i) {
        long long r = 0, c = 0;
        double w = 0.0;
        int got = pattern ? std::fscanf(f, " %lld %lld", &r, &c) : std::fscanf(f, " %lld %lld %lf", &r, &c, &w);
        if (got != (pattern ? 2 : 3)) return rtg_fail(RTG_EIO, "Could not read edge from market file");
        if (r == 0 || c == 0) return rtg_fail(RTG_EIO, "Market file is zero-indexed");
        // The reference stores such entries unchecked (bfs/graph/graph.cpp:53-54, 74-75) and
        // then indexes its CSR with them; they are rejected here.
        if (r < 0 || c < 0 || r > rows || c > cols) return rtg_fail(RTG_EIO, "Market file entry outside the matrix");
        out.push_back((int32_t)(r - 1));
        out.push_back((int32_t)(c - 1));
        if (symmetric && r != c) {
            out.push_back((int32_t)(c - 1));
            out.push_back((int32_t)(r - 1));
        }
    }
    if (rows_out) *rows_out = (int32_t)rows;
    return RTG_OK;
}

int deliver(const std::vector<int32_t> &v, int32_t *pairs, int64_t *npairs) {
    if (!npairs) return rtg_fail(RTG_EINVAL, "null npairs");
    const int64_t n = (int64_t)v.size() / 2;
    if (pairs) {
        if (*npairs < n) return rtg_fail(RTG_EINVAL, "pairs buffer too small");
        if (!v.empty()) std::memcpy(pairs, v.data(), sizeof(int32_t) * v.size());
    }
    *npairs = n;
    return RTG_OK;
}

}  // namespace

extern "C" {

RTG_API int rtg_read_binary_pairs(const char *path, int32_t *pairs, int64_t *npairs) {
    if (!path || !npairs) return rtg_fail(RTG_EINVAL, "rtg_read_binary_pairs: null argument");
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) return rtg_fail(RTG_EIO, std::string("Error: file not found: ") + path);
    const int64_t size = (int64_t)f.tellg();
    const int64_t n = size / 8;
    if (!pairs) {
        *npairs = n;
        return RTG_OK;
    }
    if (*npairs < n) return rtg_fail(RTG_EINVAL, "rtg_read_binary_pairs: buffer too small");
    f.seekg(0);
    f.read(reinterpret_cast<char *>(pairs), n * 8);
    if (!f) return rtg_fail(RTG_EIO, "rtg_read_binary_pairs: short read");
    *npairs = n;
    return RTG_OK;
}

RTG_API int rtg_read_text_pairs(const char *path, int32_t *pairs, int64_t *npairs) {
    if (!path) return rtg_fail(RTG_EINVAL, "rtg_read_text_pairs: null path");
    std::vector<int32_t> v;
    int rc = read_text_pairs(path, v);
    if (rc != RTG_OK) return rc;
    return deliver(v, pairs, npairs);
}

RTG_API int rtg_read_mtx_pairs(const char *path, int32_t *pairs, int64_t *npairs) {
    if (!path) return rtg_fail(RTG_EINVAL, "rtg_read_mtx_pairs: null path");
    std::vector<int32_t> v;
    int rc = read_mtx_pairs(path, v, nullptr);
    if (rc != RTG_OK) return rc;
    return deliver(v, pairs, npairs);
}

RTG_API int rtg_write_binary_pairs(const char *path, const int32_t *pairs, int64_t npairs) {
    if (!path || (npairs > 0 && !pairs)) return rtg_fail(RTG_EINVAL, "rtg_write_binary_pairs: bad arguments");
    std::ofstream f(path, std::ios::binary | std::ios::out);
    if (!f) return rtg_fail(RTG_EIO, std::string("Error opening file: ") + path);
    f.write(reinterpret_cast<const char *>(pairs), npairs * 8);
    if (!f) return rtg_fail(RTG_EIO, "write failed");
    return RTG_OK;
}

}  // extern "C"
