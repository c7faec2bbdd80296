// Linear-system capture files (".cfdxlin", format tag CFDX_LINEAR_REPLAY_V1):
// the reference's replay format for one FP64 CSR system with its right-hand
// side, initial guess and solver controls (core/linalg/linear_system_capture.h:
// 165-284). The text layout, the field order of the checksum and the size
// limits follow the reference so that files move both ways between it and
// this library:
//
//   CFDX_LINEAR_REPLAY_V1
//   <problem> <null space> <scaling>            (decimal)
//   <rows> <cols> <nnz> <max iterations>        (decimal)
//   <tolerance bits>                            (hex, then hex to the end)
//   <row offset>                     x rows + 1
//   <column> <value bits>            x nnz
//   <rhs bits> <initial guess bits>  x rows
//   <checksum>
//
// The checksum is 64-bit FNV-1a over the little-endian bytes of 64-bit words:
// rows, cols, nnz, problem, null space, scaling, max iterations, tolerance
// bits, every row offset, then (column, value bits) per nonzero, then
// (rhs bits, guess bits) per row.
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "cfdxhip.h"

namespace {

constexpr uint64_t kFnvOffset = 14695981039346656037ULL;
constexpr uint64_t kFnvPrime = 1099511628211ULL;
constexpr uint64_t kMaxRows = 2000000;    // replay limits of the format
constexpr uint64_t kMaxNnz = 100000000;
const char* kTag = "CFDX_LINEAR_REPLAY_V1";

struct Capture {
    cfdxhip_capture_info info{};
    std::vector<uint32_t> rowp, cols;
    std::vector<double> vals, rhs, x0;
};

thread_local std::string g_error;

inline uint64_t bits_of(double x) {
    uint64_t u;
    std::memcpy(&u, &x, 8);
    return u;
}
inline double double_of(uint64_t u) {
    double x;
    std::memcpy(&x, &u, 8);
    return x;
}

struct Fnv {
    uint64_t h = kFnvOffset;
    void word(uint64_t w) {
        for (int b = 0; b < 8; ++b) {
            h ^= (w >> (8 * b)) & 0xffu;
            h *= kFnvPrime;
        }
    }
};

uint64_t checksum(const Capture& c) {
    Fnv f;
    const uint64_t n = c.info.n;
    f.word(n);
    f.word(n);
    f.word(c.info.nnz);
    f.word(c.info.problem);
    f.word(c.info.null_space);
    f.word(c.info.scaling);
    f.word(c.info.max_iterations);
    f.word(bits_of(c.info.tolerance));
    for (uint64_t i = 0; i <= n; ++i) f.word(c.rowp[i]);
    for (uint64_t k = 0; k < c.info.nnz; ++k) {
        f.word(c.cols[k]);
        f.word(bits_of(c.vals[k]));
    }
    for (uint64_t i = 0; i < n; ++i) {
        f.word(bits_of(c.rhs[i]));
        f.word(bits_of(c.x0[i]));
    }
    return f.h;
}

// The reference's writer-side checks (linear_system_capture.h:108-137) and
// SparseMatrix::is_consistent (sparse_matrix.h:256-276).
bool validate(const Capture& c, std::string& err) {
    const auto& in = c.info;
    if (in.problem > 5 || in.null_space > 1 || in.scaling > 4) {
        err = "invalid solver metadata";
        return false;
    }
    if (in.n == 0 || in.n > kMaxRows || in.nnz > kMaxNnz) {
        err = "dimensions exceed the format's limits";
        return false;
    }
    if (in.max_iterations == 0 || in.max_iterations > 100000000 || !(in.tolerance > 0.0) ||
        !std::isfinite(in.tolerance)) {
        err = "invalid iteration controls";
        return false;
    }
    if (c.rowp.size() != in.n + 1 || c.rowp[0] != 0 || c.rowp[in.n] != in.nnz) {
        err = "invalid CSR row offsets";
        return false;
    }
    for (uint64_t i = 0; i < in.n; ++i) {
        if (c.rowp[i + 1] < c.rowp[i]) {
            err = "invalid CSR row offsets";
            return false;
        }
        for (uint32_t k = c.rowp[i]; k < c.rowp[i + 1]; ++k) {
            if (c.cols[k] >= in.n || !std::isfinite(c.vals[k]) || (k > c.rowp[i] && c.cols[k] <= c.cols[k - 1])) {
                err = "CSR columns must be in range and strictly ascending, values finite";
                return false;
            }
        }
        if (!std::isfinite(c.rhs[i]) || !std::isfinite(c.x0[i])) {
            err = "non-finite vector entry";
            return false;
        }
    }
    return true;
}

// Whitespace-separated token reader over the whole file.
class Tokens {
public:
    explicit Tokens(std::string text) : t_(std::move(text)) {}
    bool next(std::string& out) {
        while (p_ < t_.size() && std::isspace(static_cast<unsigned char>(t_[p_]))) ++p_;
        if (p_ >= t_.size()) return false;
        const size_t b = p_;
        while (p_ < t_.size() && !std::isspace(static_cast<unsigned char>(t_[p_]))) ++p_;
        out.assign(t_, b, p_ - b);
        return true;
    }
    bool number(int base, uint64_t& v) {
        std::string s;
        if (!next(s) || s.empty() || s[0] == '-' || s[0] == '+') return false;
        for (char ch : s)
            if (!(base == 16 ? std::isxdigit(static_cast<unsigned char>(ch)) : std::isdigit(static_cast<unsigned char>(ch))))
                return false;
        errno = 0;
        char* end = nullptr;
        const unsigned long long r = std::strtoull(s.c_str(), &end, base);
        if (errno != 0 || end != s.c_str() + s.size()) return false;
        v = r;
        return true;
    }

private:
    std::string t_;
    size_t p_ = 0;
};

bool read_file(const char* path, Capture& c, std::string& err) {
    std::FILE* f = std::fopen(path, "rb");
    if (!f) {
        err = "cannot open file";
        return false;
    }
    std::string text;
    char buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, got);
    std::fclose(f);
    Tokens t(std::move(text));
    std::string tag;
    if (!t.next(tag) || tag != kTag) {
        err = "unsupported format";
        return false;
    }
    uint64_t problem, nulls, scaling, rows, cols, nnz, iters;
    if (!t.number(10, problem) || !t.number(10, nulls) || !t.number(10, scaling) || !t.number(10, rows) ||
        !t.number(10, cols) || !t.number(10, nnz) || !t.number(10, iters)) {
        err = "invalid header";
        return false;
    }
    if (problem > 5 || nulls > 1 || scaling > 4) {
        err = "invalid solver metadata";
        return false;
    }
    if (rows == 0 || rows != cols || rows > kMaxRows || nnz > kMaxNnz) {
        err = "dimensions exceed supported limits";
        return false;
    }
    c.info.problem = uint32_t(problem);
    c.info.null_space = uint32_t(nulls);
    c.info.scaling = uint32_t(scaling);
    c.info.n = rows;
    c.info.nnz = nnz;
    c.info.max_iterations = iters;
    uint64_t w;
    if (!t.number(16, w)) {
        err = "truncated or invalid hexadecimal number";
        return false;
    }
    c.info.tolerance = double_of(w);
    c.rowp.resize(rows + 1);
    c.cols.resize(nnz);
    c.vals.resize(nnz);
    c.rhs.resize(rows);
    c.x0.resize(rows);
    for (auto& v : c.rowp) {
        if (!t.number(16, w) || w > 0xffffffffu) {
            err = "truncated or overflowing CSR offset";
            return false;
        }
        v = uint32_t(w);
    }
    for (uint64_t k = 0; k < nnz; ++k) {
        if (!t.number(16, w) || w >= cols) {
            err = "truncated or invalid CSR column";
            return false;
        }
        c.cols[k] = uint32_t(w);
        if (!t.number(16, w)) {
            err = "truncated or invalid hexadecimal number";
            return false;
        }
        c.vals[k] = double_of(w);
    }
    for (uint64_t i = 0; i < rows; ++i) {
        uint64_t a, b;
        if (!t.number(16, a) || !t.number(16, b)) {
            err = "truncated or invalid hexadecimal number";
            return false;
        }
        c.rhs[i] = double_of(a);
        c.x0[i] = double_of(b);
    }
    if (!t.number(16, c.info.checksum)) {
        err = "truncated or invalid hexadecimal number";
        return false;
    }
    std::string trailing;
    if (t.next(trailing)) {
        err = "unexpected trailing data";
        return false;
    }
    if (!validate(c, err)) return false;
    if (checksum(c) != c.info.checksum) {
        err = "checksum mismatch";
        return false;
    }
    return true;
}

bool write_file(const char* path, Capture& c, std::string& err) {
    if (!validate(c, err)) return false;
    c.info.checksum = checksum(c);
    const std::string tmp = std::string(path) + ".tmp";
    std::FILE* f = std::fopen(tmp.c_str(), "wb");
    if (!f) {
        err = "cannot create temporary file";
        return false;
    }
    const auto& in = c.info;
    std::fprintf(f, "%s\n%u %u %u\n%llu %llu %llu %llu\n%llx\n", kTag, in.problem, in.null_space, in.scaling,
                 (unsigned long long)in.n, (unsigned long long)in.n, (unsigned long long)in.nnz,
                 (unsigned long long)in.max_iterations, (unsigned long long)bits_of(in.tolerance));
    for (uint64_t i = 0; i <= in.n; ++i) std::fprintf(f, "%x\n", c.rowp[i]);
    for (uint64_t k = 0; k < in.nnz; ++k)
        std::fprintf(f, "%x %llx\n", c.cols[k], (unsigned long long)bits_of(c.vals[k]));
    for (uint64_t i = 0; i < in.n; ++i)
        std::fprintf(f, "%llx %llx\n", (unsigned long long)bits_of(c.rhs[i]), (unsigned long long)bits_of(c.x0[i]));
    std::fprintf(f, "%llx\n", (unsigned long long)in.checksum);
    const bool ok = std::fflush(f) == 0 && !std::ferror(f);
    const bool closed = std::fclose(f) == 0;
    if (!ok || !closed) {
        std::remove(tmp.c_str());
        err = "disk write failed";
        return false;
    }
    if (std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        err = "atomic rename failed";
        return false;
    }
    return true;
}

}  // namespace

struct cfdxhip_capture {
    Capture c;
};

extern "C" {

const char* cfdxhip_capture_last_error(void) { return g_error.c_str(); }

cfdxhip_status cfdxhip_capture_read(const char* path, cfdxhip_capture** out) {
    if (!path || !out) return CFDXHIP_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    try {
        auto h = std::make_unique<cfdxhip_capture>();
        std::string err;
        if (!read_file(path, h->c, err)) {
            g_error = "linear capture: " + err;
            return CFDXHIP_ERR_INVALID_ARGUMENT;
        }
        g_error.clear();
        *out = h.release();
        return CFDXHIP_OK;
    } catch (const std::bad_alloc&) {
        g_error = "linear capture: out of memory";
        return CFDXHIP_ERR_OUT_OF_MEMORY;
    } catch (...) {
        g_error = "linear capture: internal error";
        return CFDXHIP_ERR_INTERNAL;
    }
}

cfdxhip_status cfdxhip_capture_info_get(const cfdxhip_capture* h, cfdxhip_capture_info* info) {
    if (!h || !info) return CFDXHIP_ERR_INVALID_ARGUMENT;
    *info = h->c.info;
    return CFDXHIP_OK;
}

cfdxhip_status cfdxhip_capture_copy(const cfdxhip_capture* h, uint32_t* row_offsets, uint32_t* columns,
                                    double* values, double* rhs, double* initial_solution) {
    if (!h || !row_offsets || !columns || !values || !rhs || !initial_solution) return CFDXHIP_ERR_INVALID_ARGUMENT;
    const Capture& c = h->c;
    std::memcpy(row_offsets, c.rowp.data(), c.rowp.size() * 4);
    std::memcpy(columns, c.cols.data(), c.cols.size() * 4);
    std::memcpy(values, c.vals.data(), c.vals.size() * 8);
    std::memcpy(rhs, c.rhs.data(), c.rhs.size() * 8);
    std::memcpy(initial_solution, c.x0.data(), c.x0.size() * 8);
    return CFDXHIP_OK;
}

void cfdxhip_capture_destroy(cfdxhip_capture* h) { delete h; }

cfdxhip_status cfdxhip_capture_write(const char* path, const cfdxhip_capture_info* info, const uint32_t* row_offsets,
                                     const uint32_t* columns, const double* values, const double* rhs,
                                     const double* initial_solution, uint64_t* checksum_out) {
    if (!path || !*path || !info || !row_offsets || !rhs || !initial_solution) return CFDXHIP_ERR_INVALID_ARGUMENT;
    if (info->n == 0 || info->n > kMaxRows) {
        g_error = "linear capture: dimensions exceed the format's limits";
        return CFDXHIP_ERR_INVALID_ARGUMENT;
    }
    try {
        Capture c;
        c.info = *info;
        c.rowp.assign(row_offsets, row_offsets + info->n + 1);
        if (c.rowp[info->n] != info->nnz || info->nnz > kMaxNnz || (info->nnz && (!columns || !values))) {
            g_error = "linear capture: invalid CSR row offsets";
            return CFDXHIP_ERR_INVALID_ARGUMENT;
        }
        c.cols.assign(columns, columns + info->nnz);
        c.vals.assign(values, values + info->nnz);
        c.rhs.assign(rhs, rhs + info->n);
        c.x0.assign(initial_solution, initial_solution + info->n);
        std::string err;
        if (!write_file(path, c, err)) {
            g_error = "linear capture: " + err;
            return CFDXHIP_ERR_INVALID_ARGUMENT;
        }
        if (checksum_out) *checksum_out = c.info.checksum;
        g_error.clear();
        return CFDXHIP_OK;
    } catch (const std::bad_alloc&) {
        g_error = "linear capture: out of memory";
        return CFDXHIP_ERR_OUT_OF_MEMORY;
    } catch (...) {
        g_error = "linear capture: internal error";
        return CFDXHIP_ERR_INTERNAL;
    }
}

}  // extern "C"
