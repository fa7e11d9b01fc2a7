// validate.cpp -- parallel streaming validation of a product C = A*B, and the
// reference files it can take its expected rows from.
//
// Full mode follows validateFullStreaming (src/common/spgemm_common.cpp:720-851,
// criterion componentwise_dot_bound_v1): every output coordinate of the exact
// serial product is rebuilt in long double (reference value, rounded sum of
// |a*b|, term count), given the a-priori rounding bound of coordinateBound
// (:640-675), and the candidate is accepted where |actual - reference|,
// rounded up, is within that bound.  Partial mode follows
// buildPartialReference / comparePartialReference (:541-571, criterion
// row_nnz_exact_v1): per-row counts of the serial product's nonzero sums.
//
// Rows are independent, so the rows are validated by a pool of threads in
// fixed chunks and the per-chunk results are merged in row order: every count
// and maximum, the worst coordinate (the first, in row-major order, with the
// largest ratio) and the first eight reported differences are those of the
// serial traversal.
//
// Reference files (.spref, src/common/partial_reference.cpp:202-245; .sfull,
// src/common/full_reference.cpp:60-235) are read and written with their exact
// byte layout: little-endian integers, long doubles as (exponent tag, 64-bit
// significand limbs), a SHA-256 trailer over everything before it.  The key
// that names a file is a caller-supplied 64-hex-digit string here: the
// reference derives it from a hash of its own build configuration.
#include "spgemm_amd.h"

#include <algorithm>
#include <atomic>
#include <cfenv>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <limits>
#include <memory>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace {

using ld = long double;

// ---------------------------------------------------------------- SHA-256
// FIPS 180-4, section 6.2.
class Sha256 {
    uint32_t h_[8] = {0x6a09e667u, 0xbb67ae85u, 0x3c6ef372u, 0xa54ff53au,
                      0x510e527fu, 0x9b05688cu, 0x1f83d9abu, 0x5be0cd19u};
    unsigned char buf_[64];
    size_t used_ = 0;
    uint64_t bytes_ = 0;

    static uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }
    void block(const unsigned char *p) {
        static const uint32_t k[64] = {
            0x428a2f98u, 0x71374491u, 0xb5c0fbcfu, 0xe9b5dba5u, 0x3956c25bu, 0x59f111f1u, 0x923f82a4u, 0xab1c5ed5u,
            0xd807aa98u, 0x12835b01u, 0x243185beu, 0x550c7dc3u, 0x72be5d74u, 0x80deb1feu, 0x9bdc06a7u, 0xc19bf174u,
            0xe49b69c1u, 0xefbe4786u, 0x0fc19dc6u, 0x240ca1ccu, 0x2de92c6fu, 0x4a7484aau, 0x5cb0a9dcu, 0x76f988dau,
            0x983e5152u, 0xa831c66du, 0xb00327c8u, 0xbf597fc7u, 0xc6e00bf3u, 0xd5a79147u, 0x06ca6351u, 0x14292967u,
            0x27b70a85u, 0x2e1b2138u, 0x4d2c6dfcu, 0x53380d13u, 0x650a7354u, 0x766a0abbu, 0x81c2c92eu, 0x92722c85u,
            0xa2bfe8a1u, 0xa81a664bu, 0xc24b8b70u, 0xc76c51a3u, 0xd192e819u, 0xd6990624u, 0xf40e3585u, 0x106aa070u,
            0x19a4c116u, 0x1e376c08u, 0x2748774cu, 0x34b0bcb5u, 0x391c0cb3u, 0x4ed8aa4au, 0x5b9cca4fu, 0x682e6ff3u,
            0x748f82eeu, 0x78a5636fu, 0x84c87814u, 0x8cc70208u, 0x90befffau, 0xa4506cebu, 0xbef9a3f7u, 0xc67178f2u};
        uint32_t w[64];
        for (int t = 0; t < 16; ++t)
            w[t] = (uint32_t)p[4 * t] << 24 | (uint32_t)p[4 * t + 1] << 16 | (uint32_t)p[4 * t + 2] << 8 | p[4 * t + 3];
        for (int t = 16; t < 64; ++t) {
            const uint32_t s0 = rotr(w[t - 15], 7) ^ rotr(w[t - 15], 18) ^ (w[t - 15] >> 3);
            const uint32_t s1 = rotr(w[t - 2], 17) ^ rotr(w[t - 2], 19) ^ (w[t - 2] >> 10);
            w[t] = w[t - 16] + s0 + w[t - 7] + s1;
        }
        uint32_t a = h_[0], b = h_[1], c = h_[2], d = h_[3], e = h_[4], f = h_[5], g = h_[6], h = h_[7];
        for (int t = 0; t < 64; ++t) {
            const uint32_t t1 = h + (rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25)) + ((e & f) ^ (~e & g)) + k[t] + w[t];
            const uint32_t t2 = (rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22)) + ((a & b) ^ (a & c) ^ (b & c));
            h = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
        }
        h_[0] += a; h_[1] += b; h_[2] += c; h_[3] += d; h_[4] += e; h_[5] += f; h_[6] += g; h_[7] += h;
    }

public:
    void add(const void *data, size_t n) {
        const unsigned char *p = static_cast<const unsigned char *>(data);
        bytes_ += n;
        while (n) {
            const size_t take = std::min(n, 64 - used_);
            std::memcpy(buf_ + used_, p, take);
            used_ += take; p += take; n -= take;
            if (used_ == 64) { block(buf_); used_ = 0; }
        }
    }
    void add(const std::string &s) { add(s.data(), s.size()); }
    std::string hex() {
        const uint64_t bits = bytes_ * 8;
        const unsigned char pad = 0x80, zero = 0;
        add(&pad, 1);
        while (used_ != 56) add(&zero, 1);
        unsigned char len[8];
        for (int i = 0; i < 8; ++i) len[i] = (unsigned char)(bits >> (56 - 8 * i));
        add(len, 8);
        std::ostringstream out;
        for (uint32_t v : h_) out << std::hex << std::setw(8) << std::setfill('0') << v;
        return out.str();
    }
};

// ------------------------------------------------------------ byte codecs
void put(std::string &b, uint64_t v, int count = 8) {
    for (int i = 0; i < count; ++i) { b.push_back((char)(v & 255)); v >>= 8; }
}
uint64_t get(const std::string &b, size_t &at, int count = 8) {
    if (at + (size_t)count > b.size()) throw std::runtime_error("truncated reference record");
    uint64_t v = 0;
    for (int i = 0; i < count; ++i) v |= (uint64_t)(unsigned char)b[at++] << (8 * i);
    return v;
}
bool valid_key(const char *key) {
    if (!key || std::strlen(key) != 64) return false;
    for (const char *p = key; *p; ++p)
        if (!((*p >= '0' && *p <= '9') || (*p >= 'a' && *p <= 'f'))) return false;
    return true;
}

constexpr int kLimbs = (std::numeric_limits<ld>::digits + 63) / 64;
constexpr size_t kCellBytes = 20 + 2 * (4 + 8 * kLimbs);

// Exact long double as exponent tag (exponent + 32768, bit 16 = sign) and
// significand limbs of the frexp fraction (full_reference.cpp:60-88).
void put_real(std::string &b, ld v) {
    if (!std::isfinite(v)) throw std::runtime_error("non-finite Full reference");
    int e = 0;
    ld f = std::frexp(std::fabs(v), &e);
    if (e <= -32768 || e >= 32768) throw std::runtime_error("unsupported Full reference exponent");
    put(b, (uint64_t)(unsigned)(e + 32768) | (std::signbit(v) ? 65536u : 0u), 4);
    for (int i = 0; i < kLimbs; ++i) {
        f = std::ldexp(f, 64);
        const uint64_t digit = (uint64_t)f;
        put(b, digit);
        f -= (ld)digit;
    }
    if (f != 0) throw std::runtime_error("inexact Full reference encoding");
}
ld get_real(const std::string &b, size_t &at) {
    const uint64_t tag = get(b, at, 4);
    if (tag > 131071) throw std::runtime_error("invalid Full reference exponent");
    const int e = (int)(tag & 65535) - 32768;
    ld v = 0;
    for (int i = 0; i < kLimbs; ++i) v += std::ldexp((ld)get(b, at), e - 64 * (i + 1));
    return (tag & 65536) ? -v : v;
}

std::string sfull_header(const std::string &key, int rows, int cols) {
    if (rows < 0 || cols < 0 || std::numeric_limits<ld>::radix != 2 || kLimbs > 2)
        throw std::runtime_error("unsupported Full reference metadata");
    std::string b = "SPFR0001" + key;
    put(b, (uint64_t)rows); put(b, (uint64_t)cols);
    put(b, (uint64_t)std::numeric_limits<ld>::digits);
    put(b, (uint64_t)std::numeric_limits<ld>::max_exponent);
    put(b, kCellBytes);
    return b;
}

// ------------------------------------------------------ numeric primitives
// Directed rounding as in spgemm_common.cpp:596-609.
ld upward(ld v) { return std::nextafter(v, std::numeric_limits<ld>::infinity()); }
ld add_up(ld a, ld b) { if (a == 0) return b; if (b == 0) return a; return upward(a + b); }
ld mul_up(ld a, ld b) { return a == 0 || b == 0 ? 0 : upward(a * b); }
ld div_up(ld a, ld b) { return a == 0 ? 0 : upward(a / b); }

struct ErrModel { ld gamma, underflow; };
// gamma_(2m) and the underflow allowance for precision T (:651-661).
template <typename T>
ErrModel err_model(int64_t terms) {
    const ld ops = mul_up(2, (ld)terms);
    const ld scaled = mul_up(ops, std::numeric_limits<T>::epsilon() / 2.0L);
    if (!(scaled < 0.5L)) return {1, std::numeric_limits<ld>::infinity()};
    const ld den = std::nextafter(1 - scaled, 0.0L);
    return {div_up(scaled, den), div_up(mul_up(ops, std::numeric_limits<T>::denorm_min()), den)};
}
// coordinateBound (:663-675).
ld coordinate_bound(int64_t terms, ld reference, ld magnitude) {
    if (terms == 0) return 0;
    const ErrModel cand = err_model<double>(terms), refm = err_model<ld>(terms);
    if (!(cand.gamma < 1 && refm.gamma < 1) || !std::isfinite(reference) || !std::isfinite(magnitude))
        return std::numeric_limits<ld>::infinity();
    const ld scale = div_up(add_up(magnitude, refm.underflow), std::nextafter(1 - refm.gamma, 0.0L));
    return add_up(mul_up(add_up(cand.gamma, refm.gamma), scale), add_up(cand.underflow, refm.underflow));
}
template <typename T>
bool gradual_underflow() {
    volatile T normal = std::numeric_limits<T>::min();
    volatile T half = T(0.5);
    volatile T sub = std::numeric_limits<T>::denorm_min();
    volatile T large = std::numeric_limits<T>::max();
    return normal * half == std::numeric_limits<T>::min() / T(2) &&
           sub * large == std::numeric_limits<T>::denorm_min() * std::numeric_limits<T>::max();
}
// fullNumericEnvironment (:676-692).
bool full_environment() {
#if defined(__FAST_MATH__) || (defined(__FINITE_MATH_ONLY__) && __FINITE_MATH_ONLY__ > 0)
    return false;
#endif
    return std::numeric_limits<double>::radix == 2 && std::numeric_limits<ld>::radix == 2 &&
           std::numeric_limits<double>::digits == 53 && std::numeric_limits<ld>::digits >= 53 &&
           std::fegetround() == FE_TONEAREST && gradual_underflow<double>() && gradual_underflow<ld>();
}

// ------------------------------------------------------------- operands
struct Csr {
    int32_t rows = 0, cols = 0;
    int64_t nnz = 0;
    const int64_t *rp = nullptr;
    const int32_t *col = nullptr;
    const double *val = nullptr;
};
bool view_ok(const spg_csr_view *v) {
    return v && v->location == SPG_HOST && v->rows >= 0 && v->cols >= 0 && v->nnz >= 0 && v->row_ptr &&
           (v->nnz == 0 || (v->col_idx && v->values));
}
Csr csr(const spg_csr_view *v) { return {v->rows, v->cols, v->nnz, v->row_ptr, v->col_idx, v->values}; }

// validNumericCsr (:621-638).
bool valid_numeric(const Csr &m, bool ordered) {
    if (m.rp[0] != 0 || m.rp[m.rows] != m.nnz) return false;
    for (int32_t r = 0; r < m.rows; ++r) {
        const int64_t a = m.rp[r], b = m.rp[r + 1];
        if (a < 0 || b < a || b > m.nnz) return false;
        int prev = -1;
        for (int64_t p = a; p < b; ++p) {
            const int c = m.col[p];
            if (c < 0 || c >= m.cols || (ordered && c <= prev) || !std::isfinite(m.val[p])) return false;
            prev = c;
        }
    }
    return true;
}

struct Cell {
    int32_t column;
    double serial;
    ld reference, bound;
    int64_t terms;
};

// One row of the exact serial product with its long double reference
// (buildFullRow, :694-717): products in serial order, each coordinate summed
// from 0 in that order.  Sorting (column, serial index) gives every column's
// products in serial order.
struct RowBuilder {
    std::vector<std::pair<int32_t, int64_t>> key;  // (column, serial index)
    std::vector<double> prod;
    std::vector<ld> precise;
    void build(const Csr &A, const Csr &B, int32_t row, std::vector<Cell> &cells) {
        key.clear(); prod.clear(); precise.clear(); cells.clear();
        for (int64_t a = A.rp[row]; a < A.rp[row + 1]; ++a) {
            const int32_t k = A.col[a];
            for (int64_t b = B.rp[k]; b < B.rp[k + 1]; ++b) {
                key.emplace_back(B.col[b], (int64_t)prod.size());
                prod.push_back(A.val[a] * B.val[b]);
                precise.push_back((ld)A.val[a] * (ld)B.val[b]);
            }
        }
        std::sort(key.begin(), key.end());
        for (size_t i = 0; i < key.size();) {
            const int32_t c = key[i].first;
            double serial = 0;
            ld reference = 0, magnitude = 0;
            int64_t terms = 0;
            for (; i < key.size() && key[i].first == c; ++i) {
                const int64_t s = key[i].second;
                serial += prod[s];
                reference += precise[s];
                magnitude += std::fabs(precise[s]);
                ++terms;
            }
            cells.push_back({c, serial, reference, coordinate_bound(terms, reference, magnitude), terms});
        }
    }
};

// ----------------------------------------------------------- thread pool
template <typename F>
void parallel_chunks(int64_t n, int64_t chunk, int threads, F &&f) {
    const int64_t nchunks = (n + chunk - 1) / chunk;
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    threads = (int)std::min<int64_t>(threads, std::max<int64_t>(nchunks, 1));
    std::atomic<int64_t> next{0};
    std::exception_ptr err;
    std::mutex mu;
    auto work = [&]() {
        try {
            for (int64_t c; (c = next.fetch_add(1)) < nchunks;) f(c, c * chunk, std::min(n, (c + 1) * chunk));
        } catch (...) {
            std::lock_guard<std::mutex> g(mu);
            if (!err) err = std::current_exception();
            next.store(nchunks);
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back(work);
    work();
    for (auto &t : pool) t.join();
    if (err) std::rethrow_exception(err);
}

// ------------------------------------------------------- full comparison
struct Partial {
    double max_serial = 0;
    ld max_ref_err = 0, max_ratio = 0;
    ld worst_actual = 0, worst_reference = 0, worst_bound = 0;
    int32_t worst_row = -1, worst_col = -1;
    int64_t support = 0, rejected = 0, ref_nnz = 0, row_mismatch = 0;
    std::vector<spg_full_example> examples;  // at most 8, in traversal order
};

// The comparison of one row (the loop body of :758-826).
void compare_row(const Csr &C, int32_t row, const std::vector<Cell> &cells, Partial &p) {
    int64_t row_nnz = 0;
    for (const Cell &c : cells) row_nnz += c.serial != 0;
    p.ref_nnz += row_nnz;
    int64_t actual = C.rp[row];
    const int64_t end = C.rp[row + 1];
    if (end - actual != row_nnz) ++p.row_mismatch;
    size_t expected = 0;
    while (actual < end || expected < cells.size()) {
        Cell cell{0, 0, 0, 0, 0};
        double actual_value = 0;
        bool stored = false;
        int32_t column;
        if (expected == cells.size() || (actual < end && C.col[actual] < cells[expected].column)) {
            column = C.col[actual];
            actual_value = C.val[actual++];
            stored = true;
        } else {
            cell = cells[expected++];
            column = cell.column;
            if (actual < end && C.col[actual] == column) {
                actual_value = C.val[actual++];
                stored = true;
            }
        }
        const bool support = stored != (cell.serial != 0);
        double serial_error = std::fabs(actual_value - cell.serial);
        if (!std::isfinite(serial_error)) serial_error = std::numeric_limits<double>::infinity();
        p.max_serial = std::max(p.max_serial, serial_error);
        const ld difference = std::fabs((ld)actual_value - cell.reference);
        const ld error_upper = difference == 0 ? 0 : upward(difference);
        const ld bound = cell.bound;
        const bool accepted = std::isfinite(serial_error) && std::isfinite(error_upper) && std::isfinite(bound) &&
                              error_upper <= bound;
        const ld ratio = !std::isfinite(serial_error) || !std::isfinite(error_upper) || !std::isfinite(bound) ||
                                 (bound == 0 && difference != 0)
                             ? std::numeric_limits<ld>::infinity()
                             : bound == 0 ? 0
                             : accepted   ? std::min(1.0L, div_up(error_upper, bound))
                                          : div_up(error_upper, bound);
        if (ratio > p.max_ratio) {
            p.max_ratio = ratio;
            p.worst_row = row;
            p.worst_col = column;
            p.worst_actual = actual_value;
            p.worst_reference = cell.reference;
            p.worst_bound = bound;
        }
        p.max_ref_err = std::max(p.max_ref_err, std::isfinite(difference) ? difference : std::numeric_limits<ld>::infinity());
        if (!accepted) ++p.rejected;
        if (support) ++p.support;
        if (p.examples.size() < 8 && (support || !accepted)) {
            spg_full_example ex{};
            ex.row = row; ex.column = column; ex.accepted = accepted; ex.support_mismatch = support;
            ex.contributions = cell.terms; ex.actual = actual_value; ex.serial_expected = cell.serial;
            ex.expected = cell.reference; ex.abs_error = difference; ex.error_bound = bound; ex.error_ratio = ratio;
            p.examples.push_back(ex);
        }
    }
}

void merge(Partial &into, const Partial &p) {
    into.max_serial = std::max(into.max_serial, p.max_serial);
    into.max_ref_err = std::max(into.max_ref_err, p.max_ref_err);
    if (p.max_ratio > into.max_ratio) {  // chunks merged in row order: the first worst wins ties
        into.max_ratio = p.max_ratio;
        into.worst_row = p.worst_row; into.worst_col = p.worst_col;
        into.worst_actual = p.worst_actual; into.worst_reference = p.worst_reference; into.worst_bound = p.worst_bound;
    }
    into.support += p.support; into.rejected += p.rejected; into.ref_nnz += p.ref_nnz;
    into.row_mismatch += p.row_mismatch;
    for (const auto &e : p.examples)
        if (into.examples.size() < 8) into.examples.push_back(e);
}

void finish_report(const Partial &p, int32_t rows, spg_full_report *rep) {
    rep->rows = rows;
    rep->reference_nnz = p.ref_nnz;
    rep->row_nnz_mismatches = p.row_mismatch;
    rep->support_mismatches = p.support;
    rep->rejected_coordinates = p.rejected;
    rep->max_abs_error = p.max_serial;
    rep->max_reference_error = p.max_ref_err;
    rep->max_error_ratio = p.max_ratio;
    rep->worst_row = p.worst_row;
    rep->worst_column = p.worst_col;
    rep->worst_actual = p.worst_actual;
    rep->worst_reference = p.worst_reference;
    rep->worst_bound = p.worst_bound;
    rep->n_examples = (int32_t)p.examples.size();
    for (size_t i = 0; i < p.examples.size(); ++i) rep->examples[i] = p.examples[i];
    rep->passed = p.rejected == 0;
}

constexpr int64_t kRowChunk = 256;

// ----------------------------------------------------- .sfull streaming
struct SfullReader {
    std::ifstream file;
    Sha256 hash;
    int rows = 0, cols = 0, next = 0;
    bool finished = false;
    std::string read(size_t n, bool hashed = true) {
        std::string b(n, '\0');
        if (n && !file.read(&b[0], (std::streamsize)n)) throw std::runtime_error("missing or truncated Full reference");
        if (hashed) hash.add(b);
        return b;
    }
    SfullReader(const char *path, const std::string &key, int r, int c) : file(path, std::ios::binary), rows(r), cols(c) {
        if (!file) throw std::runtime_error(std::string("cannot open Full reference ") + path);
        const std::string expected = sfull_header(key, rows, cols);
        if (read(expected.size()) != expected) throw std::runtime_error("incompatible Full reference header");
    }
    // readRow + checkRow (full_reference.cpp:124-133, 182-201).
    void row(std::vector<Cell> &cells) {
        if (next >= rows || finished) throw std::runtime_error("Full reference row overflow");
        size_t at = 0;
        const uint64_t count = get(read(8), at);
        if (count > (uint64_t)cols) throw std::runtime_error("invalid Full reference row length");
        const std::string b = read((size_t)count * kCellBytes);
        at = 0;
        cells.clear();
        int prev = -1;
        for (uint64_t i = 0; i < count; ++i) {
            const uint64_t column = get(b, at, 4), terms = get(b, at), serial = get(b, at);
            if (column >= (uint64_t)cols || terms > (uint64_t)INT64_MAX)
                throw std::runtime_error("Full reference coordinate out of range");
            Cell c;
            c.column = (int32_t)column;
            c.terms = (int64_t)terms;
            std::memcpy(&c.serial, &serial, 8);
            c.reference = get_real(b, at);
            c.bound = get_real(b, at);
            if (c.column <= prev || c.terms <= 0 || !std::isfinite(c.serial) || !std::isfinite(c.reference) ||
                !std::isfinite(c.bound) || c.bound < 0)
                throw std::runtime_error("invalid Full reference coordinate");
            prev = c.column;
            cells.push_back(c);
        }
        ++next;
    }
    // finish (full_reference.cpp:203-208): every row read, SHA-256 trailer, no trailing data.
    void finish() {
        if (finished || next != rows) throw std::runtime_error("Full reference not completely read");
        if (read(64, false) != hash.hex() || file.peek() != std::char_traits<char>::eof() || file.bad())
            throw std::runtime_error("corrupt Full reference checksum or trailing data");
        finished = true;
    }
};

std::string encode_cells(const std::vector<Cell> &cells) {
    std::string b;
    put(b, cells.size());
    for (const Cell &c : cells) {
        put(b, (uint64_t)(uint32_t)c.column, 4);
        put(b, (uint64_t)c.terms);
        uint64_t bits;
        std::memcpy(&bits, &c.serial, 8);
        put(b, bits);
        put_real(b, c.reference);
        put_real(b, c.bound);
    }
    return b;
}

thread_local std::string t_err;

spg_status fail(spg_status s, const std::string &msg) {
    t_err = msg;
    return s;
}

}  // namespace

extern "C" {

const char *spg_validation_last_error(void) { return t_err.c_str(); }

spg_status spg_sha256_hex(const void *data, size_t n, char out[65]) {
    if ((!data && n) || !out) return fail(SPG_EINVAL, "null argument");
    Sha256 h;
    h.add(data, n);
    const std::string s = h.hex();
    std::memcpy(out, s.c_str(), 65);
    return SPG_OK;
}

spg_status spg_validate_full(const spg_csr_view *A, const spg_csr_view *B, const spg_csr_view *C,
                             const char *sfull_path, const char *key, int threads, spg_full_report *rep) {
    if (!rep) return fail(SPG_EINVAL, "null report");
    std::memset(rep, 0, sizeof(*rep));
    rep->reference_digits = std::numeric_limits<ld>::digits;
    rep->worst_row = rep->worst_column = -1;
    if (!view_ok(A) || !view_ok(B) || !view_ok(C)) return fail(SPG_EINVAL, "validation takes host CSR views");
    rep->environment_ok = full_environment();
    if (!rep->environment_ok) return SPG_OK;  // never a PASS (:730)
    const Csr a = csr(A), b = csr(B), c = csr(C);
    if (!valid_numeric(a, false) || !valid_numeric(b, false) || !valid_numeric(c, true) || a.cols != b.rows ||
        c.rows != a.rows || c.cols != b.cols) {
        rep->invalid_matrix = 1;
        return SPG_OK;
    }
    try {
        const int64_t rows = a.rows;
        const int64_t nchunks = (rows + kRowChunk - 1) / kRowChunk;
        Partial total;
        if (sfull_path) {
            if (!valid_key(key)) return fail(SPG_EINVAL, "invalid Full reference key");
            // Rows are read in order (the checksum is sequential), a block of
            // chunks at a time, and compared in parallel.
            SfullReader reader(sfull_path, key, a.rows, b.cols);
            constexpr int64_t kBlock = 64;
            std::vector<std::vector<Cell>> block;
            for (int64_t c0 = 0; c0 < nchunks; c0 += kBlock) {
                const int64_t c1 = std::min(nchunks, c0 + kBlock);
                const int64_t r0 = c0 * kRowChunk, r1 = std::min(rows, c1 * kRowChunk);
                block.resize((size_t)(r1 - r0));
                for (int64_t r = r0; r < r1; ++r) reader.row(block[(size_t)(r - r0)]);
                std::vector<Partial> parts((size_t)(c1 - c0));
                parallel_chunks(r1 - r0, kRowChunk, threads, [&](int64_t ci, int64_t lo, int64_t hi) {
                    for (int64_t r = lo; r < hi; ++r) compare_row(c, (int32_t)(r0 + r), block[(size_t)r], parts[(size_t)ci]);
                });
                for (const Partial &p : parts) merge(total, p);
            }
            reader.finish();
        } else {
            std::vector<Partial> parts((size_t)nchunks);
            parallel_chunks(rows, kRowChunk, threads, [&](int64_t ci, int64_t lo, int64_t hi) {
                RowBuilder rb;
                std::vector<Cell> cells;
                for (int64_t r = lo; r < hi; ++r) {
                    rb.build(a, b, (int32_t)r, cells);
                    compare_row(c, (int32_t)r, cells, parts[(size_t)ci]);
                }
            });
            for (const Partial &p : parts) merge(total, p);
        }
        finish_report(total, a.rows, rep);
    } catch (const std::bad_alloc &) {
        return fail(SPG_ENOMEM, "out of host memory");
    } catch (const std::exception &e) {
        return fail(SPG_EPARSE, e.what());
    }
    return SPG_OK;
}

spg_status spg_full_report_text(const spg_full_report *rep, char *buf, size_t cap, size_t *needed) {
    if (!rep) return fail(SPG_EINVAL, "null report");
    std::ostringstream o;
    o << "validation_full_strategy=streaming_roundoff_v2 validation_criterion=componentwise_dot_bound_v1"
      << " validation_reference_digits=" << rep->reference_digits
      << " validation_candidate_digits=" << std::numeric_limits<double>::digits << '\n';
    if (!rep->environment_ok) o << "validation_invalid_environment=unsupported_floating_point\n";
    else if (rep->invalid_matrix) o << "validation_invalid_matrix=1\n";
    else {
        o << std::setprecision(std::numeric_limits<ld>::max_digits10);
        for (int i = 0; i < rep->n_examples; ++i) {
            const spg_full_example &e = rep->examples[i];
            o << "validation_difference row=" << e.row << " column=" << e.column << " actual=" << e.actual
              << " expected=" << e.expected << " serial_expected=" << e.serial_expected << " abs_error=" << e.abs_error
              << " error_bound=" << e.error_bound << " error_ratio=" << e.error_ratio
              << " contributions=" << e.contributions << " accepted=" << e.accepted
              << " support_mismatch=" << e.support_mismatch << '\n';
        }
        o << "validation_index_base=0 validation_full_rows=" << rep->rows
          << " validation_reference_nnz=" << rep->reference_nnz
          << " validation_row_nnz_mismatches=" << rep->row_nnz_mismatches
          << " validation_support_mismatches=" << rep->support_mismatches
          << " validation_rejected_coordinates=" << rep->rejected_coordinates
          << " validation_max_abs_error=" << rep->max_abs_error
          << " validation_max_reference_error=" << rep->max_reference_error
          << " validation_max_error_ratio=" << rep->max_error_ratio
          << " validation_worst_ratio_row=" << rep->worst_row << " validation_worst_ratio_column=" << rep->worst_column
          << " validation_worst_ratio_actual=" << rep->worst_actual
          << " validation_worst_ratio_reference=" << rep->worst_reference
          << " validation_worst_ratio_bound=" << rep->worst_bound << '\n';
    }
    const std::string s = o.str();
    if (needed) *needed = s.size() + 1;
    if (buf && cap) {
        const size_t n = std::min(cap - 1, s.size());
        std::memcpy(buf, s.data(), n);
        buf[n] = '\0';
    }
    return SPG_OK;
}

// buildPartialReference (:541-554): rows of the serial product (first product
// initialises, later ones add in order) and their nonzero counts.
spg_status spg_partial_counts(const spg_csr_view *A, const spg_csr_view *B, int threads, int64_t *row_nnz,
                              int32_t *finite) {
    if (!view_ok(A) || !view_ok(B) || !row_nnz || !finite) return fail(SPG_EINVAL, "validation takes host CSR views");
    const Csr a = csr(A), b = csr(B);
    if (a.cols != b.rows) return fail(SPG_EINVAL, "SpGEMM requires A.cols == B.rows");
    for (int64_t e = 0; e < a.nnz; ++e)
        if (a.col[e] < 0 || a.col[e] >= b.rows) return fail(SPG_ERANGE, "A contains a column index outside B's row range");
    std::atomic<int> all_finite{1};
    try {
        parallel_chunks(a.rows, kRowChunk, threads, [&](int64_t, int64_t lo, int64_t hi) {
            std::vector<std::pair<int32_t, int64_t>> key;
            std::vector<double> prod;
            bool fin = true;
            for (int64_t r = lo; r < hi; ++r) {
                key.clear(); prod.clear();
                for (int64_t x = a.rp[r]; x < a.rp[r + 1]; ++x)
                    for (int64_t y = b.rp[a.col[x]]; y < b.rp[a.col[x] + 1]; ++y) {
                        key.emplace_back(b.col[y], (int64_t)prod.size());
                        prod.push_back(a.val[x] * b.val[y]);
                    }
                std::sort(key.begin(), key.end());
                int64_t count = 0;
                for (size_t i = 0; i < key.size();) {
                    const int32_t c = key[i].first;
                    double v = prod[key[i].second];
                    for (++i; i < key.size() && key[i].first == c; ++i) v += prod[key[i].second];
                    if (!std::isfinite(v)) fin = false;
                    count += v != 0.0;
                }
                row_nnz[r] = count;
            }
            if (!fin) all_finite.store(0);
        });
    } catch (const std::bad_alloc &) {
        return fail(SPG_ENOMEM, "out of host memory");
    }
    *finite = all_finite.load();
    return SPG_OK;
}

// comparePartialReference (:556-571).
spg_status spg_validate_partial(const spg_csr_view *A, const spg_csr_view *B, const spg_csr_view *C,
                                const char *spref_path, const char *key, int threads, int32_t *passed,
                                int64_t *row_nnz_mismatches) {
    if (!view_ok(A) || !view_ok(B) || !view_ok(C) || !passed || !row_nnz_mismatches)
        return fail(SPG_EINVAL, "validation takes host CSR views");
    *passed = 0;
    *row_nnz_mismatches = 0;
    if (A->cols != B->rows) return fail(SPG_EINVAL, "SpGEMM requires A.cols == B.rows");
    if (C->rows != A->rows || C->cols != B->cols) return SPG_OK;
    std::vector<int64_t> ref((size_t)A->rows);
    int32_t finite = 1;
    spg_status s = spref_path ? spg_spref_read(spref_path, key, A->rows, B->cols, ref.data(), &finite)
                              : spg_partial_counts(A, B, threads, ref.data(), &finite);
    if (s != SPG_OK) return s;
    bool ok = finite != 0;
    for (int64_t e = 0; e < C->nnz; ++e) ok = ok && std::isfinite(C->values[e]);
    int64_t mism = 0;
    for (int32_t r = 0; r < A->rows; ++r) mism += ref[(size_t)r] != C->row_ptr[r + 1] - C->row_ptr[r];
    *row_nnz_mismatches = mism;
    *passed = ok && mism == 0;
    return SPG_OK;
}

// encodePartialReference / decodePartialReference (partial_reference.cpp:202-233).
spg_status spg_spref_write(const char *path, const char *key, int32_t rows, int32_t cols, const int64_t *row_nnz,
                           int32_t finite) {
    if (!path || !valid_key(key) || rows < 0 || cols < 0 || (rows && !row_nnz))
        return fail(SPG_EINVAL, "invalid partial reference arguments");
    std::string b = std::string("SPPR0001") + key;
    put(b, (uint64_t)rows); put(b, (uint64_t)cols); put(b, finite ? 1 : 0);
    for (int32_t r = 0; r < rows; ++r) {
        if (row_nnz[r] < 0) return fail(SPG_EINVAL, "negative reference count");
        put(b, (uint64_t)row_nnz[r]);
    }
    Sha256 h;
    h.add(b);
    b += h.hex();
    const std::string tmp = std::string(path) + ".tmp";
    {
        std::ofstream f(tmp, std::ios::binary | std::ios::trunc);
        if (!f.write(b.data(), (std::streamsize)b.size())) return fail(SPG_EPARSE, "cannot write partial reference");
    }
    if (std::rename(tmp.c_str(), path) != 0) return fail(SPG_EPARSE, "cannot publish partial reference");
    return SPG_OK;
}

spg_status spg_spref_read(const char *path, const char *key, int32_t rows, int32_t cols, int64_t *row_nnz,
                          int32_t *finite) {
    if (!path || !valid_key(key) || rows < 0 || cols < 0 || (rows && !row_nnz) || !finite)
        return fail(SPG_EINVAL, "invalid partial reference arguments");
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    const uint64_t expected = 160ull + 8ull * (uint64_t)rows;
    if (!f || (uint64_t)f.tellg() != expected)
        return fail(SPG_EPARSE, std::string("missing or incomplete partial reference: ") + path);
    std::string b((size_t)expected, '\0');
    f.seekg(0);
    if (!f.read(&b[0], (std::streamsize)b.size())) return fail(SPG_EPARSE, "cannot read partial reference");
    Sha256 h;
    h.add(b.data(), b.size() - 64);
    if (b.compare(0, 8, "SPPR0001") != 0 || b.compare(8, 64, key) != 0 || h.hex() != b.substr(b.size() - 64))
        return fail(SPG_EPARSE, "incomplete, corrupt or incompatible partial reference");
    size_t at = 72;
    const uint64_t r = get(b, at), c = get(b, at), fin = get(b, at);
    if (r != (uint64_t)rows || c != (uint64_t)cols || fin > 1) return fail(SPG_EPARSE, "invalid partial reference metadata");
    for (int32_t i = 0; i < rows; ++i) {
        const uint64_t v = get(b, at);
        if (v > (uint64_t)INT64_MAX) return fail(SPG_EPARSE, "partial reference count exceeds signed 64-bit range");
        row_nnz[i] = (int64_t)v;
    }
    *finite = (int32_t)fin;
    return SPG_OK;
}

// prepareFullReferenceFile + the numeric row builder (full_reference.cpp:210-235,
// spgemm_common.cpp:852-861): rows built in parallel, written in order, then
// the SHA-256 trailer; published by rename.
spg_status spg_sfull_write(const spg_csr_view *A, const spg_csr_view *B, const char *path, const char *key,
                           int threads) {
    if (!view_ok(A) || !view_ok(B) || !path || !valid_key(key)) return fail(SPG_EINVAL, "invalid Full reference arguments");
    if (!full_environment()) return fail(SPG_EDEVICE, "unsupported floating-point environment");
    const Csr a = csr(A), b = csr(B);
    if (!valid_numeric(a, false) || !valid_numeric(b, false) || a.cols != b.rows)
        return fail(SPG_EINVAL, "invalid Full reference inputs");
    const std::string tmp = std::string(path) + ".tmp";
    try {
        std::ofstream f(tmp, std::ios::binary | std::ios::trunc);
        if (!f) return fail(SPG_EPARSE, "cannot create Full reference temporary file");
        Sha256 h;
        auto write = [&](const std::string &s) {
            if (!f.write(s.data(), (std::streamsize)s.size())) throw std::runtime_error("cannot write Full reference");
            h.add(s);
        };
        write(sfull_header(key, a.rows, b.cols));
        const int64_t rows = a.rows;
        constexpr int64_t kBlockRows = 64 * kRowChunk;
        std::vector<std::string> enc;
        for (int64_t r0 = 0; r0 < rows; r0 += kBlockRows) {
            const int64_t r1 = std::min(rows, r0 + kBlockRows);
            enc.assign((size_t)(r1 - r0), std::string());
            parallel_chunks(r1 - r0, kRowChunk, threads, [&](int64_t, int64_t lo, int64_t hi) {
                RowBuilder rb;
                std::vector<Cell> cells;
                for (int64_t r = lo; r < hi; ++r) {
                    rb.build(a, b, (int32_t)(r0 + r), cells);
                    enc[(size_t)r] = encode_cells(cells);
                }
            });
            for (const std::string &s : enc) write(s);
        }
        const std::string trailer = h.hex();
        if (!f.write(trailer.data(), 64)) throw std::runtime_error("cannot write Full reference");
        f.close();
        if (!f) throw std::runtime_error("cannot write Full reference");
    } catch (const std::bad_alloc &) {
        std::remove(tmp.c_str());
        return fail(SPG_ENOMEM, "out of host memory");
    } catch (const std::exception &e) {
        std::remove(tmp.c_str());
        return fail(SPG_EPARSE, e.what());
    }
    if (std::rename(tmp.c_str(), path) != 0) return fail(SPG_EPARSE, "cannot publish Full reference");
    return SPG_OK;
}

// Streaming reader of a .sfull file (FullReferenceReader).
struct spg_sfull {
    std::unique_ptr<SfullReader> r;
    std::vector<Cell> cells;
};

spg_status spg_sfull_open(const char *path, const char *key, int32_t rows, int32_t cols, spg_sfull **out) {
    if (!path || !valid_key(key) || !out) return fail(SPG_EINVAL, "invalid Full reference arguments");
    *out = nullptr;
    try {
        auto s = std::make_unique<spg_sfull>();
        s->r = std::make_unique<SfullReader>(path, key, rows, cols);
        *out = s.release();
    } catch (const std::exception &e) {
        return fail(SPG_EPARSE, e.what());
    }
    return SPG_OK;
}

spg_status spg_sfull_next_row(spg_sfull *s, int64_t *ncells) {
    if (!s || !ncells) return fail(SPG_EINVAL, "null argument");
    try {
        s->r->row(s->cells);
    } catch (const std::exception &e) {
        return fail(SPG_EPARSE, e.what());
    }
    *ncells = (int64_t)s->cells.size();
    return SPG_OK;
}

spg_status spg_sfull_row_cells(const spg_sfull *s, int32_t *column, double *serial, long double *reference,
                               long double *bound, int64_t *terms) {
    if (!s) return fail(SPG_EINVAL, "null argument");
    for (size_t i = 0; i < s->cells.size(); ++i) {
        const Cell &c = s->cells[i];
        if (column) column[i] = c.column;
        if (serial) serial[i] = c.serial;
        if (reference) reference[i] = c.reference;
        if (bound) bound[i] = c.bound;
        if (terms) terms[i] = c.terms;
    }
    return SPG_OK;
}

spg_status spg_sfull_finish(spg_sfull *s) {
    if (!s) return fail(SPG_EINVAL, "null argument");
    try {
        s->r->finish();
    } catch (const std::exception &e) {
        return fail(SPG_EPARSE, e.what());
    }
    return SPG_OK;
}

void spg_sfull_close(spg_sfull *s) { delete s; }

}  // extern "C"
