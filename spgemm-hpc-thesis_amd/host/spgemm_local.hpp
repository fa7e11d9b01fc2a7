// spgemm_local.hpp -- C++17 face of libspgemm_amd.so in the reference's own terms.
//
// Mirrors the reference's local-multiply interface so a caller written against
// it switches by changing a namespace:
//   CsrMatrix            src/common/csr_matrix.hpp:43-49
//   RowBlock             src/common/spgemm_common.hpp:56-59
//   spgemm(...)          src/baselines/spgemm_exchange.hpp:37-39 (body spgemm_exchange.cpp:258-325)
//   LocalAccumulator     src/trident/common/matrix_blocks.hpp:23-32 (body matrix_blocks.cpp:152-200)
//   readMatrixMarket     src/common/matrix_market.cpp:28-144
// Error behaviour follows the reference: spgemm() compares no shapes; an A
// column that names neither a local B row nor a halo slot (at -np 1: a column
// outside B's rows) raises MissingRowError, which a caller holding an MPI
// communicator turns into the reference's fail() (spgemm_exchange.cpp:82-85).
// std::invalid_argument reports malformed operands and the accumulator's
// mismatched shapes, std::overflow_error count overflow.
//
// Header-only; link with -lspgemm_amd.
#pragma once

#include "spgemm_amd.h"

#include <cstdint>
#include <limits>
#include <new>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace spgemm_amd {

using CsrOffset = std::int64_t;

struct CsrMatrix {
    int rows = 0;
    int cols = 0;
    std::vector<CsrOffset> rowPtr;
    std::vector<int> columnIndices;
    std::vector<double> values;
};

struct RowBlock {
    int firstRow = 0;
    int rows = 0;
};

// The gathered remote B rows in slot order: slot s is global B row
// globalRows[s] (RemoteSparseRows, spgemm_exchange.hpp:8-20, with
// plan.rowToSlot flattened into globalRows).
struct HaloRows {
    std::vector<int> globalRows;
    std::vector<CsrOffset> rowPtr{0};
    std::vector<int> columnIndices;
    std::vector<double> values;
};

struct MissingRowError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

namespace detail {

[[noreturn]] inline void raise(spg_status s, const std::string &msg) {
    switch (s) {
    case SPG_EINVAL: throw std::invalid_argument(msg);
    case SPG_EOVERFLOW: throw std::overflow_error(msg);
    case SPG_ENOMEM: throw std::bad_alloc();
    case SPG_EMISSING_ROW: throw MissingRowError(msg);
    default: throw std::runtime_error(msg);
    }
}

inline void check(spg_status s, const spg_context *ctx) {
    if (s == SPG_OK) return;
    std::string msg = ctx ? spg_last_error(ctx) : "";
    if (msg.empty()) msg = spg_status_string(s);
    raise(s, msg);
}

inline spg_csr_view view(const CsrMatrix &m) {
    spg_csr_view v{};
    v.rows = m.rows;
    v.cols = m.cols;
    v.nnz = m.rowPtr.empty() ? 0 : m.rowPtr.back();
    v.row_ptr = m.rowPtr.data();
    v.col_idx = m.columnIndices.data();
    v.values = m.values.data();
    v.location = SPG_HOST;
    return v;
}

inline std::size_t count(std::int64_t n) {
    if (n < 0) throw std::overflow_error("CSR allocation is not addressable: " + std::to_string(n));
    return static_cast<std::size_t>(n);
}

}  // namespace detail

// One device context (HIP stream, workspace and result buffers).  The
// reference's spgemm() is a free function; the engine keeps its device
// workspace between calls, so callers hold one Engine per host thread.
class Engine {
public:
    explicit Engine(int device = 0) { detail::check(spg_create(&ctx_, device), nullptr); }
    ~Engine() {
        if (ctx_) spg_destroy(ctx_);
    }
    Engine(const Engine &) = delete;
    Engine &operator=(const Engine &) = delete;
    Engine(Engine &&o) noexcept : ctx_(std::exchange(o.ctx_, nullptr)) {}
    spg_context *get() const { return ctx_; }
    spg_stats lastStats() const {
        spg_stats s{};
        detail::check(spg_last_stats(ctx_, &s), ctx_);
        return s;
    }

private:
    spg_context *ctx_ = nullptr;
};

// The local multiply: rows of A times B, where global B row k is local row
// k - localBBlock.firstRow when it lies in localBBlock, else the halo slot
// holding it.  With no halo and localBBlock.firstRow == 0 this is the -np 1
// product.  Result: C.rows = A.rows, C.cols = B.cols, columns ascending,
// exact zeros dropped, values bit-identical to the reference's serial order.
inline CsrMatrix spgemm(Engine &engine, const CsrMatrix &localMatrixA, RowBlock localBBlock,
                        const CsrMatrix &localMatrixB, const HaloRows *remoteBRows = nullptr) {
    spg_context *ctx = engine.get();
    const spg_csr_view a = detail::view(localMatrixA);
    const spg_csr_view b = detail::view(localMatrixB);
    spg_halo_view h{};
    if (remoteBRows) {
        h.n_rows = static_cast<int32_t>(remoteBRows->globalRows.size());
        h.global_rows = remoteBRows->globalRows.data();
        h.row_ptr = remoteBRows->rowPtr.data();
        h.col_idx = remoteBRows->columnIndices.data();
        h.values = remoteBRows->values.data();
        h.location = SPG_HOST;
    }
    if (localBBlock.rows != localMatrixB.rows)
        throw std::invalid_argument("localBBlock.rows must equal localMatrixB.rows");
    std::int64_t nnz = 0;
    detail::check(spg_multiply_begin(ctx, &a, &b, localBBlock.firstRow, remoteBRows ? &h : nullptr, &nnz), ctx);
    CsrMatrix c;
    c.rows = localMatrixA.rows;
    c.cols = localMatrixB.cols;
    c.rowPtr.resize(detail::count(static_cast<std::int64_t>(c.rows) + 1));
    c.columnIndices.resize(detail::count(nnz));
    c.values.resize(detail::count(nnz));
    detail::check(spg_multiply_finish(ctx, c.rowPtr.data(), c.columnIndices.data(), c.values.data()), ctx);
    return c;
}

// Staged accumulation: acc[col] += a*b over successive addProduct calls (stage
// order, then A entries, then B entries in stored order); finish() drops exact
// zeros and returns columns ascending.
class LocalAccumulator {
public:
    LocalAccumulator(Engine &engine, int rows, int cols) : engine_(&engine), rows_(rows), cols_(cols) {
        detail::check(spg_accumulator_create(engine.get(), rows, cols, &acc_), engine.get());
    }
    ~LocalAccumulator() {
        if (acc_) spg_accumulator_destroy(acc_);
    }
    LocalAccumulator(const LocalAccumulator &) = delete;
    LocalAccumulator &operator=(const LocalAccumulator &) = delete;

    void addProduct(const CsrMatrix &a, const CsrMatrix &b) {
        const spg_csr_view va = detail::view(a), vb = detail::view(b);
        detail::check(spg_accumulator_add(acc_, &va, &vb), engine_->get());
    }

    CsrMatrix finish() const {
        std::int64_t nnz = 0;
        detail::check(spg_accumulator_finish(acc_, &nnz), engine_->get());
        CsrMatrix c;
        c.rows = rows_;
        c.cols = cols_;
        c.rowPtr.resize(detail::count(static_cast<std::int64_t>(c.rows) + 1));
        c.columnIndices.resize(detail::count(nnz));
        c.values.resize(detail::count(nnz));
        detail::check(spg_accumulator_copy(acc_, c.rowPtr.data(), c.columnIndices.data(), c.values.data()),
                      engine_->get());
        return c;
    }

private:
    Engine *engine_;
    int rows_, cols_;
    spg_accumulator *acc_ = nullptr;
};

// MatrixMarket coordinate file -> CSR in file order (duplicates, explicit
// zeros and the symmetric mirror entries kept where the reference keeps them).
inline CsrMatrix readMatrixMarket(const std::string &path) {
    spg_matrix *m = nullptr;
    const spg_status s = spg_matrix_read_mm(path.c_str(), &m);
    if (s != SPG_OK) detail::raise(s, spg_matrix_last_error());
    spg_csr_view v{};
    spg_matrix_view(m, &v);
    CsrMatrix c;
    c.rows = v.rows;
    c.cols = v.cols;
    c.rowPtr.assign(v.row_ptr, v.row_ptr + v.rows + 1);
    c.columnIndices.assign(v.col_idx, v.col_idx + v.nnz);
    c.values.assign(v.values, v.values + v.nnz);
    spg_matrix_free(m);
    return c;
}

// validateProduct (src/common/spgemm_common.cpp:865-904), multithreaded on
// the host.  The expected rows come from a reference file (`referenceFile`,
// .spref for Partial, .sfull for Full, named by `key`) or are rebuilt from A
// and B when it is empty.  `text` holds the validation_* lines of Full mode.
enum class ValidationMode { None, Partial, Full };
struct ValidationResult {
    ValidationMode mode = ValidationMode::None;
    bool passed = true;
    double maxAbsoluteError = 0;
    std::int64_t rowNnzMismatches = 0;
    long double maxErrorRatio = 0;
    std::int64_t rejectedCoordinates = 0;
    int referenceDigits = 0;
    std::string text;
};

inline ValidationResult validateProduct(const CsrMatrix &A, const CsrMatrix &B, const CsrMatrix &C,
                                        ValidationMode mode, const std::string &referenceFile = "",
                                        const std::string &key = "", int threads = 0) {
    ValidationResult r;
    r.mode = mode;
    if (mode == ValidationMode::None) return r;
    const spg_csr_view a = detail::view(A), b = detail::view(B), c = detail::view(C);
    const char *file = referenceFile.empty() ? nullptr : referenceFile.c_str();
    if (mode == ValidationMode::Partial) {
        std::int32_t passed = 0;
        const spg_status s = spg_validate_partial(&a, &b, &c, file, key.c_str(), threads, &passed, &r.rowNnzMismatches);
        if (s != SPG_OK) detail::raise(s, spg_validation_last_error());
        r.passed = passed != 0;
        return r;
    }
    spg_full_report rep;
    const spg_status s = spg_validate_full(&a, &b, &c, file, key.c_str(), threads, &rep);
    if (s != SPG_OK) detail::raise(s, spg_validation_last_error());
    std::size_t need = 0;
    spg_full_report_text(&rep, nullptr, 0, &need);
    r.text.assign(need, '\0');
    spg_full_report_text(&rep, &r.text[0], need, nullptr);
    r.text.resize(need ? need - 1 : 0);
    r.passed = rep.environment_ok && !rep.invalid_matrix && rep.passed;
    r.maxAbsoluteError = rep.environment_ok && !rep.invalid_matrix ? rep.max_abs_error
                                                                   : std::numeric_limits<double>::infinity();
    r.rowNnzMismatches = rep.row_nnz_mismatches;
    r.maxErrorRatio = rep.max_error_ratio;
    r.rejectedCoordinates = rep.rejected_coordinates;
    r.referenceDigits = rep.reference_digits;
    return r;
}

}  // namespace spgemm_amd
