// host_matrix.cpp -- host CSR matrices: the MatrixMarket reader of the C ABI.
//
// Follows readMatrixMarket (src/common/matrix_market.cpp:28-144):
//  * banner "%%MatrixMarket matrix coordinate <field> <symmetry>" with nothing
//    after it; field real/integer/pattern, symmetry general/symmetric/
//    skew-symmetric/hermitian (case-insensitive words after the banner);
//  * comment ('%') and blank lines skipped before the size line and between
//    entries; the size line is "rows cols entries" with rows, cols > 0;
//  * entries are 1-based, checked against the shape; 'D'/'d' exponents accepted;
//    values must be finite; nothing may follow the last token of a line;
//  * non-general storage needs a square matrix and emits the mirrored entry
//    right after each off-diagonal one (negated for skew-symmetric, whose
//    diagonal must be zero);
//  * anything other than comments after the declared entries is an error;
//  * the CSR keeps file order inside each row, duplicates and explicit zeros
//    (stable counting sort by row).
// The file is read once; the entry lines are cut into line-aligned chunks
// that threads count, then parse (integers and values with from_chars, strtod
// for a value token from_chars does not take whole), and a stable counting
// sort over the chunks in file order builds the CSR.  Errors are those of a
// sequential read: the first bad declared entry in file order, then too few
// or too many entry lines.
//
// Binary CSR ("SPGCSR01", int32 rows and cols, int64 nnz, row_ptr, col, val,
// little endian): a matrix parsed once reloads at disk speed.
#include "spgemm_amd.h"

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <new>
#include <string>
#include <thread>
#include <vector>

struct spg_matrix {
    int32_t rows = 0;
    int32_t cols = 0;
    std::vector<int64_t> row_ptr;
    std::vector<int32_t> col;
    std::vector<double> val;
};

namespace {

thread_local std::string g_matrix_error;

spg_status fail(spg_status s, const std::string &msg) {
    g_matrix_error = msg;
    return s;
}

std::vector<std::string> tokens(const std::string &line) {
    std::vector<std::string> out;
    size_t i = 0;
    while (i < line.size()) {
        while (i < line.size() && std::isspace(static_cast<unsigned char>(line[i]))) ++i;
        size_t j = i;
        while (j < line.size() && !std::isspace(static_cast<unsigned char>(line[j]))) ++j;
        if (j > i) out.emplace_back(line.substr(i, j - i));
        i = j;
    }
    return out;
}

std::string lower(std::string s) {
    for (char &c : s) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
    return s;
}

bool to_i64(const std::string &t, long long &v) {
    if (t.empty()) return false;
    errno = 0;
    char *end = nullptr;
    v = std::strtoll(t.c_str(), &end, 10);
    return end && *end == '\0' && errno != ERANGE;
}

bool to_i32(const std::string &t, int32_t &v) {
    long long x;
    if (!to_i64(t, x) || x < std::numeric_limits<int32_t>::min() || x > std::numeric_limits<int32_t>::max())
        return false;
    v = static_cast<int32_t>(x);
    return true;
}

struct Entry {
    int32_t row, col;
    double val;
};

bool is_space(char c) { return std::isspace(static_cast<unsigned char>(c)) != 0; }

// A line [b, e) without its '\n': blank or comment as blank_or_comment().
bool blank_or_comment(const char *b, const char *e) {
    for (; b < e; ++b) {
        const char c = *b;
        if (c == ' ' || c == '\t' || c == '\r' || c == '\n') continue;
        return c == '%';
    }
    return true;
}

// Integer token as strtoll base 10 would take it whole (optional sign).
bool tok_i32(const char *b, const char *e, int32_t &v) {
    if (b < e && *b == '+') {
        ++b;
        if (b < e && (*b == '+' || *b == '-')) return false;
    }
    if (b == e) return false;
    long long x = 0;
    const auto r = std::from_chars(b, e, x);
    if (r.ec != std::errc() || r.ptr != e) return false;
    if (x < std::numeric_limits<int32_t>::min() || x > std::numeric_limits<int32_t>::max()) return false;
    v = static_cast<int32_t>(x);
    return true;
}

// Value token: the characters to_f64 accepts ('D'/'d' exponents mapped to
// 'E'), then from_chars; a token it does not take whole goes to strtod.  Both
// round correctly, so the value is strtod's either way.
bool tok_f64(const char *b, const char *e, double &v) {
    if (b == e) return false;
    char buf[128];
    const size_t n = static_cast<size_t>(e - b);
    std::string big;
    char *t = buf;
    if (n >= sizeof buf) {
        big.assign(b, e);
        t = &big[0];
    } else {
        std::memcpy(buf, b, n);
        buf[n] = '\0';
    }
    for (size_t i = 0; i < n; ++i) {
        char &c = t[i];
        if (c == 'D') c = 'E';
        else if (c == 'd') c = 'e';
        if (!(std::isdigit(static_cast<unsigned char>(c)) || c == '+' || c == '-' || c == '.' || c == 'e' || c == 'E'))
            return false;
    }
    const auto r = std::from_chars(t, t + n, v);
    if (r.ec == std::errc() && r.ptr == t + n) return true;
    char *end = nullptr;
    v = std::strtod(t, &end);
    return end && *end == '\0';
}

// Up to `max` whitespace-separated tokens of [b, e); returns the count (max+1
// when there are more).
int split(const char *b, const char *e, const char **tb, const char **te, int max) {
    int n = 0;
    while (b < e) {
        while (b < e && is_space(*b)) ++b;
        if (b == e) break;
        const char *s = b;
        while (b < e && !is_space(*b)) ++b;
        if (n == max) return max + 1;
        tb[n] = s;
        te[n] = b;
        ++n;
    }
    return n;
}

// Next line [b, e) of the buffer at p (e excludes '\n'); false at the end.
bool next_line(const char *&p, const char *end, const char *&b, const char *&e) {
    if (p >= end) return false;
    b = p;
    const char *nl = static_cast<const char *>(std::memchr(p, '\n', static_cast<size_t>(end - p)));
    e = nl ? nl : end;
    p = nl ? nl + 1 : end;
    return true;
}

bool next_content(const char *&p, const char *end, std::string &line) {
    const char *b, *e;
    while (next_line(p, end, b, e)) {
        if (!blank_or_comment(b, e)) {
            line.assign(b, e);
            return true;
        }
    }
    return false;
}

struct ChunkOut {
    std::vector<Entry> ents;
    int64_t first_bad = -1;  // global content-line index of the first bad declared entry
    spg_status bad_status = SPG_OK;
    std::string bad_msg;
    int64_t lines = 0;
};

int reader_threads(size_t bytes) {
    unsigned hw = std::thread::hardware_concurrency();
    if (hw == 0) hw = 1;
    const size_t by_size = bytes / (size_t(4) << 20) + 1;  // at least 4 MB per thread
    size_t t = std::min<size_t>(std::min<size_t>(hw, 64), by_size);
    return static_cast<int>(t < 1 ? 1 : t);
}

template <typename F>
void run_threads(int n, F &&f) {
    std::vector<std::thread> pool;
    for (int t = 1; t < n; ++t) pool.emplace_back(f, t);
    f(0);
    for (auto &th : pool) th.join();
}

spg_status read_mm(const char *path, spg_matrix &m) {
    std::FILE *f = std::fopen(path, "rb");
    if (!f) return fail(SPG_EPARSE, std::string("cannot open Matrix Market file: ") + path);
    std::vector<char> buf;
    {
        struct Closer {
            std::FILE *f;
            ~Closer() { std::fclose(f); }
        } closer{f};
        if (std::fseek(f, 0, SEEK_END) == 0) {
            const long sz = std::ftell(f);
            if (sz > 0) buf.resize(static_cast<size_t>(sz));
            std::fseek(f, 0, SEEK_SET);
        }
        size_t got = buf.empty() ? 0 : std::fread(buf.data(), 1, buf.size(), f);
        buf.resize(got);
        char tmp[65536];  // streams whose size is unknown
        for (size_t n; (n = std::fread(tmp, 1, sizeof tmp, f)) > 0;) buf.insert(buf.end(), tmp, tmp + n);
    }
    const char *p = buf.data(), *end = buf.data() + buf.size();
    std::string line;
    {
        const char *b, *e;
        if (!next_line(p, end, b, e)) return fail(SPG_EPARSE, std::string("Matrix Market file is empty: ") + path);
        line.assign(b, e);
    }
    std::vector<std::string> h = tokens(line);
    h.resize(h.size() < 5 ? 5 : h.size());
    const std::string field = lower(h[3]), symmetry = lower(h[4]);
    if (h[0] != "%%MatrixMarket" || lower(h[1]) != "matrix" || lower(h[2]) != "coordinate" || h.size() > 5)
        return fail(SPG_EPARSE, std::string("only Matrix Market coordinate matrices are supported: ") + path);
    const bool pattern = field == "pattern";
    if (field != "real" && field != "integer" && !pattern)
        return fail(SPG_EPARSE, std::string("only real, integer, and pattern Matrix Market fields are supported: ") + path);
    const bool general = symmetry == "general", skew = symmetry == "skew-symmetric";
    if (!general && symmetry != "symmetric" && !skew && symmetry != "hermitian")
        return fail(SPG_EPARSE, "unsupported Matrix Market symmetry: " + symmetry);
    if (!next_content(p, end, line)) return fail(SPG_EPARSE, std::string("Matrix Market file has no size line: ") + path);
    std::vector<std::string> d = tokens(line);
    int32_t rows = 0, cols = 0;
    long long entries = 0;
    if (d.size() != 3 || !to_i32(d[0], rows) || !to_i32(d[1], cols) || !to_i64(d[2], entries) || rows <= 0 ||
        cols <= 0 || entries < 0)
        return fail(SPG_EPARSE, std::string("invalid Matrix Market dimensions: ") + path);
    const bool mirrored = !general;
    if (mirrored && rows != cols)
        return fail(SPG_EPARSE, "Matrix Market " + symmetry + " storage requires a square matrix: " + path);

    // Entry region in line-aligned chunks, one per thread.
    const int T = reader_threads(static_cast<size_t>(end - p));
    std::vector<const char *> cut(static_cast<size_t>(T) + 1);
    cut[0] = p;
    cut[static_cast<size_t>(T)] = end;
    for (int t = 1; t < T; ++t) {
        const char *c = p + (end - p) * t / T;
        if (c < cut[static_cast<size_t>(t) - 1]) c = cut[static_cast<size_t>(t) - 1];
        const char *nl = c < end ? static_cast<const char *>(std::memchr(c, '\n', static_cast<size_t>(end - c))) : nullptr;
        cut[static_cast<size_t>(t)] = nl ? nl + 1 : end;
    }
    // Pass 1: content lines per chunk.
    std::vector<ChunkOut> ch(static_cast<size_t>(T));
    run_threads(T, [&](int t) {
        const char *q = cut[static_cast<size_t>(t)], *qe = cut[static_cast<size_t>(t) + 1];
        const char *b, *e;
        int64_t n = 0;
        while (next_line(q, qe, b, e)) n += !blank_or_comment(b, e);
        ch[static_cast<size_t>(t)].lines = n;
    });
    std::vector<int64_t> first_line(static_cast<size_t>(T) + 1, 0);
    for (int t = 0; t < T; ++t) first_line[static_cast<size_t>(t) + 1] = first_line[static_cast<size_t>(t)] + ch[static_cast<size_t>(t)].lines;
    const int64_t total_lines = first_line[static_cast<size_t>(T)];
    // Pass 2: parse the declared entries (content lines with index < entries).
    run_threads(T, [&](int t) {
        ChunkOut &o = ch[static_cast<size_t>(t)];
        const int64_t g0 = first_line[static_cast<size_t>(t)];
        const int64_t take = std::max<int64_t>(0, std::min<int64_t>(o.lines, entries - g0));
        o.ents.reserve(static_cast<size_t>(mirrored ? 2 * take : take));
        const char *q = cut[static_cast<size_t>(t)], *qe = cut[static_cast<size_t>(t) + 1];
        const char *b, *e;
        int64_t k = 0;
        const char *tb[4], *te[4];
        const int want = pattern ? 2 : 3;
        while (k < take && next_line(q, qe, b, e)) {
            if (blank_or_comment(b, e)) continue;
            const int n = split(b, e, tb, te, 3);
            int32_t r = 0, c = 0;
            double v = 1.0;
            if (n != want || !tok_i32(tb[0], te[0], r) || !tok_i32(tb[1], te[1], c) ||
                (!pattern && !tok_f64(tb[2], te[2], v)) || !std::isfinite(v)) {
                o.first_bad = g0 + k;
                o.bad_status = SPG_EPARSE;
                o.bad_msg = std::string("invalid Matrix Market entry: ") + path;
                return;
            }
            if (r < 1 || r > rows || c < 1 || c > cols) {
                o.first_bad = g0 + k;
                o.bad_status = SPG_EPARSE;
                o.bad_msg = std::string("Matrix Market entry index is out of range: ") + path;
                return;
            }
            --r;
            --c;
            if (skew && r == c && v != 0.0) {
                o.first_bad = g0 + k;
                o.bad_status = SPG_EPARSE;
                o.bad_msg = std::string("skew-symmetric diagonal must be zero: ") + path;
                return;
            }
            o.ents.push_back({r, c, v});
            if (mirrored && r != c) o.ents.push_back({c, r, skew ? -v : v});
            ++k;
        }
    });
    for (const ChunkOut &o : ch)
        if (o.first_bad >= 0) return fail(o.bad_status, o.bad_msg);  // chunks in file order: the first bad line
    if (total_lines < entries) return fail(SPG_EPARSE, std::string("unexpected end of Matrix Market entries: ") + path);
    if (total_lines > entries) return fail(SPG_EPARSE, std::string("more Matrix Market entries than declared: ") + path);

    // CSR by a stable counting sort: chunk t's entries of a row follow those
    // of chunks < t, each chunk in file order.
    m.rows = rows;
    m.cols = cols;
    const size_t R = static_cast<size_t>(rows);
    std::vector<std::vector<int64_t>> cnt(static_cast<size_t>(T));
    run_threads(T, [&](int t) {
        auto &c = cnt[static_cast<size_t>(t)];
        c.assign(R, 0);
        for (const Entry &x : ch[static_cast<size_t>(t)].ents) ++c[static_cast<size_t>(x.row)];
    });
    m.row_ptr.assign(R + 1, 0);
    for (size_t r = 0; r < R; ++r) {
        int64_t run = m.row_ptr[r];
        for (int t = 0; t < T; ++t) {
            const int64_t n = cnt[static_cast<size_t>(t)][r];
            cnt[static_cast<size_t>(t)][r] = run;  // chunk t's first slot in row r
            run += n;
        }
        m.row_ptr[r + 1] = run;
    }
    const size_t nnz = static_cast<size_t>(m.row_ptr[R]);
    m.col.resize(nnz);
    m.val.resize(nnz);
    run_threads(T, [&](int t) {
        auto &cur = cnt[static_cast<size_t>(t)];
        for (const Entry &x : ch[static_cast<size_t>(t)].ents) {
            const int64_t q = cur[static_cast<size_t>(x.row)]++;
            m.col[static_cast<size_t>(q)] = x.col;
            m.val[static_cast<size_t>(q)] = x.val;
        }
    });
    return SPG_OK;
}

}  // namespace

extern "C" {

const char *spg_matrix_last_error(void) { return g_matrix_error.c_str(); }

spg_status spg_matrix_read_mm(const char *path, spg_matrix **out) {
    if (!path || !out) return fail(SPG_EINVAL, "null argument");
    *out = nullptr;
    spg_matrix *m = new (std::nothrow) spg_matrix();
    if (!m) return fail(SPG_ENOMEM, "out of memory");
    spg_status s;
    try {
        s = read_mm(path, *m);
    } catch (const std::bad_alloc &) {
        s = fail(SPG_ENOMEM, "out of memory");
    }
    if (s != SPG_OK) {
        delete m;
        return s;
    }
    *out = m;
    return SPG_OK;
}

spg_status spg_matrix_view(const spg_matrix *m, spg_csr_view *out) {
    if (!m || !out) return fail(SPG_EINVAL, "null argument");
    out->rows = m->rows;
    out->cols = m->cols;
    out->nnz = static_cast<int64_t>(m->val.size());
    out->row_ptr = m->row_ptr.data();
    out->col_idx = m->col.data();
    out->values = m->val.data();
    out->location = SPG_HOST;
    return SPG_OK;
}

spg_status spg_matrix_free(spg_matrix *m) {
    delete m;
    return SPG_OK;
}

spg_status spg_matrix_write_csr(const spg_csr_view *m, const char *path) {
    if (!m || !path || m->location != SPG_HOST || m->rows < 0 || m->cols < 0 || m->nnz < 0 || !m->row_ptr ||
        (m->nnz && (!m->col_idx || !m->values)))
        return fail(SPG_EINVAL, "binary CSR takes a host view");
    const std::string tmp = std::string(path) + ".tmp";
    std::FILE *f = std::fopen(tmp.c_str(), "wb");
    if (!f) return fail(SPG_EPARSE, std::string("cannot create ") + tmp);
    const char magic[8] = {'S', 'P', 'G', 'C', 'S', 'R', '0', '1'};
    const int32_t dims[2] = {m->rows, m->cols};
    const size_t R = static_cast<size_t>(m->rows) + 1, N = static_cast<size_t>(m->nnz);
    bool ok = std::fwrite(magic, 1, 8, f) == 8 && std::fwrite(dims, 4, 2, f) == 2 && std::fwrite(&m->nnz, 8, 1, f) == 1 &&
              std::fwrite(m->row_ptr, 8, R, f) == R && (!N || (std::fwrite(m->col_idx, 4, N, f) == N &&
                                                              std::fwrite(m->values, 8, N, f) == N));
    ok = (std::fclose(f) == 0) && ok;
    if (!ok || std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        return fail(SPG_EPARSE, std::string("cannot write binary CSR ") + path);
    }
    return SPG_OK;
}

spg_status spg_matrix_read_csr(const char *path, spg_matrix **out) {
    if (!path || !out) return fail(SPG_EINVAL, "null argument");
    *out = nullptr;
    std::FILE *f = std::fopen(path, "rb");
    if (!f) return fail(SPG_EPARSE, std::string("cannot open binary CSR ") + path);
    struct Closer {
        std::FILE *f;
        ~Closer() { std::fclose(f); }
    } closer{f};
    char magic[8];
    int32_t dims[2];
    int64_t nnz = 0;
    if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, "SPGCSR01", 8) != 0 || std::fread(dims, 4, 2, f) != 2 ||
        std::fread(&nnz, 8, 1, f) != 1 || dims[0] < 0 || dims[1] < 0 || nnz < 0)
        return fail(SPG_EPARSE, std::string("not a binary CSR file: ") + path);
    spg_matrix *m = new (std::nothrow) spg_matrix();
    if (!m) return fail(SPG_ENOMEM, "out of memory");
    try {
        m->rows = dims[0];
        m->cols = dims[1];
        const size_t R = static_cast<size_t>(dims[0]) + 1, N = static_cast<size_t>(nnz);
        m->row_ptr.resize(R);
        m->col.resize(N);
        m->val.resize(N);
        bool ok = std::fread(m->row_ptr.data(), 8, R, f) == R && (!N || (std::fread(m->col.data(), 4, N, f) == N &&
                                                                       std::fread(m->val.data(), 8, N, f) == N));
        ok = ok && std::fgetc(f) == EOF && m->row_ptr[0] == 0 && m->row_ptr[R - 1] == nnz;
        for (size_t r = 1; ok && r < R; ++r) ok = m->row_ptr[r] >= m->row_ptr[r - 1];
        for (size_t e = 0; ok && e < N; ++e) ok = m->col[e] >= 0 && m->col[e] < dims[1];
        if (!ok) {
            delete m;
            return fail(SPG_EPARSE, std::string("corrupt or truncated binary CSR: ") + path);
        }
    } catch (const std::bad_alloc &) {
        delete m;
        return fail(SPG_ENOMEM, "out of memory");
    }
    *out = m;
    return SPG_OK;
}

}  // extern "C"
