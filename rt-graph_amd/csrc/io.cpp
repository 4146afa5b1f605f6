// Host-side edge-list ingest in the reference's three on-disk formats.
//
//   binary  little-endian int32 (src, dst) pairs, npairs = file_size / 8
//           (Graph::ReadBinaryEdges, tc/graph/graph.cpp:23-39)
//   text    one "src dst" pair per line, unparsable lines skipped with a message
//           (Graph::Txt2Bin, tc/graph/graph.cpp:179-214); parsed in place from one read of
//           the file, as the reference's `int` extraction reads each line
//   .mtx    Matrix Market coordinate pattern/real/integer, 1-based, converted to 0-based;
//           "symmetric" files get every off-diagonal entry twice, (r,c) then (c,r)
//           (Graph::ReadMtx, bfs/graph/graph.cpp:17-122)
//
// All readers use the two-call convention: pairs == NULL returns the pair count only (the text
// reader reports unparsable lines in the second, filling call).
#include "rtg_internal.hpp"

#include <cctype>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

namespace {

// The whole file, NUL-terminated (the parsers below scan it in place).
int slurp(const char *path, std::vector<char> &buf, const char *open_error) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return rtg_fail(RTG_EIO, std::string(open_error) + path);
    struct Closer {
        FILE *f;
        ~Closer() { std::fclose(f); }
    } closer{f};
    if (std::fseek(f, 0, SEEK_END) != 0) return rtg_fail(RTG_EIO, std::string("cannot seek ") + path);
    const long size = std::ftell(f);
    if (size < 0) return rtg_fail(RTG_EIO, std::string("cannot size ") + path);
    std::rewind(f);
    buf.resize((size_t)size + 1);
    if (size > 0 && std::fread(buf.data(), 1, (size_t)size, f) != (size_t)size)
        return rtg_fail(RTG_EIO, std::string("short read of ") + path);
    buf[(size_t)size] = '\0';
    buf.resize((size_t)size + 1);
    return RTG_OK;
}

inline bool space_in_line(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }
inline bool space_any(char c) { return space_in_line(c) || c == '\n'; }

// One int32 as `std::istream >> int32_t` reads it: leading white space skipped, an optional sign,
// decimal digits up to the first other character; no digit, or a value outside int32, fails.
inline bool parse_i32(const char *&p, const char *end, int32_t &v) {
    while (p < end && space_in_line(*p)) ++p;
    bool neg = false;
    if (p < end && (*p == '+' || *p == '-')) neg = *p++ == '-';
    if (p >= end || *p < '0' || *p > '9') return false;
    int64_t x = 0;
    bool over = false;
    for (; p < end && *p >= '0' && *p <= '9'; ++p) {
        x = x * 10 + (*p - '0');
        if (x > (int64_t)INT32_MAX + 1) {
            over = true;
            x = (int64_t)INT32_MAX + 1;
        }
    }
    if (neg) x = -x;
    if (over || x < INT32_MIN || x > INT32_MAX) return false;
    v = (int32_t)x;
    return true;
}

// One integer as scanf's %lld reads it (white space including newlines skipped, optional sign,
// decimal digits); values beyond +-2^62 saturate (such entries are out of any matrix anyway).
inline bool parse_i64(const char *&p, long long &v) {
    while (space_any(*p)) ++p;
    bool neg = false;
    if (*p == '+' || *p == '-') neg = *p++ == '-';
    if (*p < '0' || *p > '9') return false;
    long long x = 0;
    for (; *p >= '0' && *p <= '9'; ++p) x = x < (1ll << 62) / 10 ? x * 10 + (*p - '0') : (1ll << 62);
    v = neg ? -x : x;
    return true;
}

// Text edge list: every line that begins with two int32 values gives a pair (anything after them
// is ignored); other lines are skipped, with a message unless empty or a '#' / '%' comment.
int read_text_pairs(const char *path, std::vector<int32_t> &out, bool report) {
    std::vector<char> buf;
    if (const int rc = slurp(path, buf, "cannot open ")) return rc;
    const char *p = buf.data(), *end = buf.data() + buf.size() - 1;
    out.reserve((size_t)(end - p) / 6);
    while (p < end) {
        const char *eol = static_cast<const char *>(std::memchr(p, '\n', (size_t)(end - p)));
        if (!eol) eol = end;
        const char *q = p;
        int32_t a, b;
        if (parse_i32(q, eol, a) && parse_i32(q, eol, b)) {
            out.push_back(a);
            out.push_back(b);
        } else if (report && eol > p && *p != '#' && *p != '%') {
            std::fprintf(stderr, "Error parsing line: %.*s\n", (int)(eol - p), p);
        }
        p = eol + 1;
    }
    return RTG_OK;
}

int read_mtx_pairs(const char *path, std::vector<int32_t> &out, int32_t *rows_out) {
    std::vector<char> buf;
    if (const int rc = slurp(path, buf, "File could not be opened: ")) return rc;
    const char *p = buf.data(), *end = buf.data() + buf.size() - 1;
    auto next_line = [&](std::string &line) {  // false at the end of the file
        if (p >= end) return false;
        const char *eol = static_cast<const char *>(std::memchr(p, '\n', (size_t)(end - p)));
        if (!eol) eol = end;
        line.assign(p, eol);
        p = eol < end ? eol + 1 : end;
        return true;
    };
    std::string banner, line;
    if (!next_line(banner)) return rtg_fail(RTG_EIO, "Could not process Matrix Market banner");
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
    while (next_line(line)) {
        if (!line.empty() && line[0] == '%') continue;
        if (std::sscanf(line.c_str(), "%lld %lld %lld", &rows, &cols, &nnz) == 3) {
            have_size = true;
            break;
        }
        return rtg_fail(RTG_EIO, "Could not read file info (M, N, NNZ)");
    }
    if (!have_size || rows < 0 || cols < 0 || nnz < 0 || rows > INT32_MAX || cols > INT32_MAX)
        return rtg_fail(RTG_EIO, "Could not read file info (M, N, NNZ)");
    out.reserve((size_t)std::min<long long>(nnz, (long long)(end - p) / 4 + 1) * (symmetric ? 4 : 2));
    for (long long i = 0; i < nnz; ++i) {
        long long r = 0, c = 0;
        bool ok = parse_i64(p, r) && parse_i64(p, c);
        if (ok && !pattern) {  // the value column is read and ignored
            char *after = nullptr;
            (void)std::strtod(p, &after);
            ok = after != p;
            p = after;
        }
        if (!ok) return rtg_fail(RTG_EIO, "Could not read edge from market file");
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
    int rc = read_text_pairs(path, v, pairs != nullptr);  // unparsable lines reported once, by the filling call
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
