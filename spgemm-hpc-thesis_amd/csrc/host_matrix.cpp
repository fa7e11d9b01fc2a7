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
#include "spgemm_amd.h"

#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <new>
#include <string>
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

bool blank_or_comment(const std::string &line) {
    for (char c : line) {
        if (c == ' ' || c == '\t' || c == '\r' || c == '\n') continue;
        return c == '%';
    }
    return true;
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

bool to_f64(const std::string &t, double &v) {
    if (t.empty()) return false;
    for (char c : t)
        if (!(std::isdigit(static_cast<unsigned char>(c)) || c == '+' || c == '-' || c == '.' || c == 'e' || c == 'E'))
            return false;
    char *end = nullptr;
    v = std::strtod(t.c_str(), &end);
    return end && *end == '\0';
}

bool next_content_line(std::FILE *f, std::string &line) {
    for (;;) {
        line.clear();
        int ch;
        bool any = false;
        while ((ch = std::fgetc(f)) != EOF) {
            any = true;
            if (ch == '\n') break;
            line.push_back(static_cast<char>(ch));
        }
        if (!any) return false;
        if (!blank_or_comment(line)) return true;
    }
}

struct Entry {
    int32_t row, col;
    double val;
};

spg_status read_mm(const char *path, spg_matrix &m) {
    std::FILE *f = std::fopen(path, "r");
    if (!f) return fail(SPG_EPARSE, std::string("cannot open Matrix Market file: ") + path);
    struct Closer {
        std::FILE *f;
        ~Closer() { std::fclose(f); }
    } closer{f};
    std::string line;
    {
        int ch;
        bool any = false;
        while ((ch = std::fgetc(f)) != EOF) {
            any = true;
            if (ch == '\n') break;
            line.push_back(static_cast<char>(ch));
        }
        if (!any) return fail(SPG_EPARSE, std::string("Matrix Market file is empty: ") + path);
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
    if (!next_content_line(f, line)) return fail(SPG_EPARSE, std::string("Matrix Market file has no size line: ") + path);
    std::vector<std::string> d = tokens(line);
    int32_t rows = 0, cols = 0;
    long long entries = 0;
    if (d.size() != 3 || !to_i32(d[0], rows) || !to_i32(d[1], cols) || !to_i64(d[2], entries) || rows <= 0 ||
        cols <= 0 || entries < 0)
        return fail(SPG_EPARSE, std::string("invalid Matrix Market dimensions: ") + path);
    const bool mirrored = !general;
    if (mirrored && rows != cols)
        return fail(SPG_EPARSE, "Matrix Market " + symmetry + " storage requires a square matrix: " + path);
    std::vector<Entry> ents;
    try {
        ents.reserve(static_cast<size_t>(mirrored ? 2 * entries : entries));
    } catch (const std::exception &) {
        return fail(SPG_ENOMEM, "Matrix Market reserve failed");
    }
    for (long long i = 0; i < entries; ++i) {
        if (!next_content_line(f, line)) return fail(SPG_EPARSE, std::string("unexpected end of Matrix Market entries: ") + path);
        for (char &c : line) {
            if (c == 'D') c = 'E';
            else if (c == 'd') c = 'e';
        }
        std::vector<std::string> t = tokens(line);
        int32_t r = 0, c = 0;
        double v = 1.0;
        const size_t want = pattern ? 2 : 3;
        if (t.size() != want || !to_i32(t[0], r) || !to_i32(t[1], c) || (!pattern && !to_f64(t[2], v)) ||
            !std::isfinite(v))
            return fail(SPG_EPARSE, std::string("invalid Matrix Market entry: ") + path);
        if (r < 1 || r > rows || c < 1 || c > cols)
            return fail(SPG_EPARSE, std::string("Matrix Market entry index is out of range: ") + path);
        --r;
        --c;
        if (skew && r == c && v != 0.0) return fail(SPG_EPARSE, std::string("skew-symmetric diagonal must be zero: ") + path);
        ents.push_back({r, c, v});
        if (mirrored && r != c) ents.push_back({c, r, skew ? -v : v});
    }
    while (next_content_line(f, line))
        return fail(SPG_EPARSE, std::string("more Matrix Market entries than declared: ") + path);
    m.rows = rows;
    m.cols = cols;
    m.row_ptr.assign(static_cast<size_t>(rows) + 1, 0);
    for (const Entry &e : ents) ++m.row_ptr[static_cast<size_t>(e.row) + 1];
    for (int32_t r = 0; r < rows; ++r) m.row_ptr[static_cast<size_t>(r) + 1] += m.row_ptr[static_cast<size_t>(r)];
    m.col.resize(ents.size());
    m.val.resize(ents.size());
    std::vector<int64_t> cursor(m.row_ptr.begin(), m.row_ptr.end() - 1);
    for (const Entry &e : ents) {
        const int64_t p = cursor[static_cast<size_t>(e.row)]++;
        m.col[static_cast<size_t>(p)] = e.col;
        m.val[static_cast<size_t>(p)] = e.val;
    }
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

}  // extern "C"
