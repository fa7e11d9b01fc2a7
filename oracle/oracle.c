/*
 * oracle.c -- CPU restatement of the reference's local SpGEMM path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls this
 * file: it is the checker used by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py.  The product path (spgemm-hpc-thesis_amd/) runs
 * the HIP kernels and fails loudly when they are missing.
 *
 * Every function below restates one piece of malchioman/SpGEMM-hpc-thesis and
 * cites the lines it follows.  Pins (tests/test_oracle.py):
 *   orc_product         - reference binary outputs (oracle/_ref) on every fixture,
 *                         dense brute-force oracle on small-integer cases, known
 *                         answers of tests/spgemm_serial_smoke.cpp.
 *   orc_partial_counts  - .spref-equivalent counts from oracle/_ref
 *                         (buildPartialReference), known answers of
 *                         tests/partial_reference_test.cpp:62,82-89.
 *   orc_full_validate   - known answers of tests/spgemm_full_streaming.cpp
 *                         (HV15R witnesses with pinned max_abs_error, order case,
 *                         corruption/underflow/subnormal cases) and
 *                         tests/spgemm_validation_smoke.cpp.
 *   orc_scalar_products - closed form (nnzA * per-row B length) and
 *                         tests/spgemm_serial_smoke.cpp:76-80.
 *   orc_mm_read         - oracle/_ref readMatrixMarket on every fixture file and
 *                         tests/matrix_market_contract.cpp expectations.
 *   orc_acc_*           - oracle/_ref trident::LocalAccumulator on staged products.
 *   orc_percentile90    - hand-computed nearest-rank cases.
 *
 * Arithmetic: compiled with -ffp-contract=off and without -march (so no FMA
 * contraction), like the reference's -O3 -DNDEBUG build.
 */
#define _GNU_SOURCE
#include <ctype.h>
#include <errno.h>
#include <fenv.h>
#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    int32_t rows;
    int32_t cols;
    int64_t nnz;
    int64_t *row_ptr;
    int32_t *col;
    double *val;
} orc_csr;

enum {
    ORC_OK = 0,
    ORC_EINVAL = -1,    /* A.cols != B.rows, malformed CSR */
    ORC_ERANGE = -2,    /* A column outside B's row range */
    ORC_EOVERFLOW = -3, /* int64 count overflow */
    ORC_ENOMEM = -4,
    ORC_EPARSE = -5     /* MatrixMarket format error */
};

static _Thread_local char g_err[512];

const char *orc_last_error(void) { return g_err; }

static int fail(int code, const char *fmt, const char *arg) {
    snprintf(g_err, sizeof g_err, fmt, arg ? arg : "");
    return code;
}

void orc_free(orc_csr *m) {
    if (!m) return;
    free(m->row_ptr);
    free(m->col);
    free(m->val);
    m->row_ptr = NULL;
    m->col = NULL;
    m->val = NULL;
}

static int checked_add(int64_t a, int64_t b, int64_t *out) {
    /* csr_matrix.hpp:21-27 (checkedCsrAdd) */
    if (a < 0 || b < 0 || a > INT64_MAX - b) return 0;
    *out = a + b;
    return 1;
}

static int cmp_i32(const void *x, const void *y) {
    int32_t a = *(const int32_t *)x, b = *(const int32_t *)y;
    return (a > b) - (a < b);
}

/* ------------------------------------------------------------------------ */
/* scalarMultiplicationCount: spgemm_common.cpp:474-487                       */
/* ------------------------------------------------------------------------ */
int orc_scalar_products(const orc_csr *A, const orc_csr *B, int64_t *out) {
    if (A->cols != B->rows) return fail(ORC_EINVAL, "SpGEMM requires A.cols == B.rows%s", "");
    int64_t count = 0;
    for (int32_t r = 0; r < A->rows; ++r) {
        for (int64_t e = A->row_ptr[r]; e < A->row_ptr[r + 1]; ++e) {
            int32_t k = A->col[e];
            if (k < 0 || k >= B->rows)
                return fail(ORC_ERANGE, "A contains a column index outside B's row range%s", "");
            if (!checked_add(count, B->row_ptr[k + 1] - B->row_ptr[k], &count))
                return fail(ORC_EOVERFLOW, "scalar products: overflowing count%s", "");
        }
    }
    *out = count;
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* Serial row-wise product.                                                  */
/*   accumulation: spgemm_common.cpp:131-156 (forEachSerialProductRow) and    */
/*                 spgemm_exchange.cpp:274-295 (same loop inside spgemm()):   */
/*     for A entries in stored order, for B entries in stored order:          */
/*       p = a*b; first occurrence of a column stores p, later ones add p.    */
/*   output:       spgemm_common.cpp:158-178 / spgemm_exchange.cpp:297-324:   */
/*     ascending columns, entries with value == 0.0 dropped, int64 prefix.    */
/* The hash map is replaced by a dense sparse accumulator (value + marker +   */
/* touched list); iteration order of the map never reaches the output        */
/* because the columns are sorted, so the results are identical.             */
/* Rows are split in contiguous static blocks across threads so the output   */
/* assembly is a concatenation in row order (row results do not depend on    */
/* the thread count).                                                        */
/* ------------------------------------------------------------------------ */
typedef struct {
    int64_t n, cap;
    int32_t *col;
    double *val;
} obuf;

static int obuf_push(obuf *b, int32_t c, double v) {
    if (b->n == b->cap) {
        int64_t nc = b->cap ? 2 * b->cap : 1024;
        int32_t *c2 = realloc(b->col, (size_t)nc * sizeof *c2);
        if (!c2) return 0;
        b->col = c2;
        double *v2 = realloc(b->val, (size_t)nc * sizeof *v2);
        if (!v2) return 0;
        b->val = v2;
        b->cap = nc;
    }
    b->col[b->n] = c;
    b->val[b->n] = v;
    b->n++;
    return 1;
}

int orc_product(const orc_csr *A, const orc_csr *B, orc_csr *C, int threads) {
    if (A->cols != B->rows) return fail(ORC_EINVAL, "SpGEMM requires A.cols == B.rows%s", "");
    memset(C, 0, sizeof *C);
    C->rows = A->rows;
    C->cols = B->cols;
    const int32_t rows = A->rows;
    int64_t *row_nnz = calloc((size_t)rows + 1, sizeof(int64_t));
    if (!row_nnz) return fail(ORC_ENOMEM, "out of memory%s", "");
    int nt = 1;
#ifdef _OPENMP
    nt = threads > 0 ? threads : omp_get_max_threads();
#else
    (void)threads;
#endif
    if (nt < 1) nt = 1;
    obuf *bufs = calloc((size_t)nt, sizeof(obuf));
    int status = ORC_OK;
    if (!bufs) {
        free(row_nnz);
        return fail(ORC_ENOMEM, "out of memory%s", "");
    }
#ifdef _OPENMP
#pragma omp parallel num_threads(nt)
#endif
    {
        int t = 0, T = 1;
#ifdef _OPENMP
        t = omp_get_thread_num();
        T = omp_get_num_threads();
#endif
        int32_t r0 = (int32_t)(((int64_t)rows * t) / T);
        int32_t r1 = (int32_t)(((int64_t)rows * (t + 1)) / T);
        size_t ncols = B->cols > 0 ? (size_t)B->cols : 1;
        double *acc = malloc(ncols * sizeof(double));
        int32_t *mark = malloc(ncols * sizeof(int32_t));
        int32_t *touched = malloc(ncols * sizeof(int32_t));
        int local = ORC_OK;
        if (!acc || !mark || !touched) local = ORC_ENOMEM;
        if (local == ORC_OK) {
            for (size_t j = 0; j < ncols; ++j) mark[j] = -1;
            for (int32_t r = r0; r < r1 && local == ORC_OK; ++r) {
                int64_t nt_ = 0;
                for (int64_t e = A->row_ptr[r]; e < A->row_ptr[r + 1]; ++e) {
                    const int32_t k = A->col[e];
                    if (k < 0 || k >= B->rows) {
                        local = ORC_ERANGE;
                        break;
                    }
                    const double a = A->val[e];
                    for (int64_t q = B->row_ptr[k]; q < B->row_ptr[k + 1]; ++q) {
                        const int32_t j = B->col[q];
                        const double p = a * B->val[q];
                        if (mark[j] != r) {
                            mark[j] = r;
                            acc[j] = p;
                            touched[nt_++] = j;
                        } else {
                            acc[j] += p;
                        }
                    }
                }
                if (local != ORC_OK) break;
                qsort(touched, (size_t)nt_, sizeof(int32_t), cmp_i32);
                int64_t kept = 0;
                for (int64_t i = 0; i < nt_; ++i) {
                    const int32_t j = touched[i];
                    if (acc[j] != 0.0) {
                        if (!obuf_push(&bufs[t], j, acc[j])) {
                            local = ORC_ENOMEM;
                            break;
                        }
                        ++kept;
                    }
                }
                row_nnz[r + 1] = kept;
            }
        }
        free(acc);
        free(mark);
        free(touched);
        if (local != ORC_OK) {
#ifdef _OPENMP
#pragma omp critical
#endif
            status = local;
        }
    }
    if (status == ORC_OK) {
        for (int32_t r = 0; r < rows; ++r) {
            if (!checked_add(row_nnz[r], row_nnz[r + 1], &row_nnz[r + 1])) {
                status = ORC_EOVERFLOW;
                break;
            }
        }
    }
    if (status == ORC_OK) {
        const int64_t nnz = row_nnz[rows];
        C->nnz = nnz;
        C->row_ptr = row_nnz;
        C->col = malloc((size_t)(nnz ? nnz : 1) * sizeof(int32_t));
        C->val = malloc((size_t)(nnz ? nnz : 1) * sizeof(double));
        if (!C->col || !C->val) status = ORC_ENOMEM;
        int64_t pos = 0;
        for (int t = 0; t < nt && status == ORC_OK; ++t) {
            if (bufs[t].n) {
                memcpy(C->col + pos, bufs[t].col, (size_t)bufs[t].n * sizeof(int32_t));
                memcpy(C->val + pos, bufs[t].val, (size_t)bufs[t].n * sizeof(double));
            }
            pos += bufs[t].n;
        }
        row_nnz = NULL;
    }
    for (int t = 0; t < nt; ++t) {
        free(bufs[t].col);
        free(bufs[t].val);
    }
    free(bufs);
    free(row_nnz);
    if (status != ORC_OK) {
        orc_free(C);
        if (status == ORC_ERANGE) return fail(status, "A contains a column index outside B's row range%s", "");
        if (status == ORC_EOVERFLOW) return fail(status, "local result prefix overflow%s", "");
        return fail(status, "out of memory%s", "");
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* buildPartialReference: spgemm_common.cpp:541-554 -- same ordered traversal */
/* as the serial product; counts accumulated values != 0.0 per row; the       */
/* reference is "finite" iff every accumulated value is finite.              */
/* ------------------------------------------------------------------------ */
int orc_partial_counts(const orc_csr *A, const orc_csr *B, int64_t *row_nnz, int32_t *finite) {
    if (A->cols != B->rows) return fail(ORC_EINVAL, "SpGEMM requires A.cols == B.rows%s", "");
    size_t ncols = B->cols > 0 ? (size_t)B->cols : 1;
    double *acc = malloc(ncols * sizeof(double));
    int32_t *mark = malloc(ncols * sizeof(int32_t));
    int32_t *touched = malloc(ncols * sizeof(int32_t));
    if (!acc || !mark || !touched) {
        free(acc); free(mark); free(touched);
        return fail(ORC_ENOMEM, "out of memory%s", "");
    }
    for (size_t j = 0; j < ncols; ++j) mark[j] = -1;
    *finite = 1;
    for (int32_t r = 0; r < A->rows; ++r) {
        int64_t nt_ = 0;
        for (int64_t e = A->row_ptr[r]; e < A->row_ptr[r + 1]; ++e) {
            const int32_t k = A->col[e];
            if (k < 0 || k >= B->rows) {
                free(acc); free(mark); free(touched);
                return fail(ORC_ERANGE, "A contains a column index outside B's row range%s", "");
            }
            const double a = A->val[e];
            for (int64_t q = B->row_ptr[k]; q < B->row_ptr[k + 1]; ++q) {
                const int32_t j = B->col[q];
                const double p = a * B->val[q];
                if (mark[j] != r) {
                    mark[j] = r;
                    acc[j] = p;
                    touched[nt_++] = j;
                } else {
                    acc[j] += p;
                }
            }
        }
        int64_t count = 0;
        for (int64_t i = 0; i < nt_; ++i) {
            const double v = acc[touched[i]];
            if (!isfinite(v)) *finite = 0;
            if (v != 0.0) ++count;
        }
        row_nnz[r] = count;
    }
    free(acc); free(mark); free(touched);
    return ORC_OK;
}

/* comparePartialReference: spgemm_common.cpp:556-572.  Returns 1 on
 * PARTIAL_PASS.  rows/cols mismatch -> fail with 0 mismatches counted. */
int orc_partial_compare(int32_t ref_rows, int32_t ref_cols, int32_t ref_finite, const int64_t *ref_nnz,
                        const orc_csr *C, int64_t *mismatches) {
    *mismatches = 0;
    if (C->rows != ref_rows || C->cols != ref_cols) return 0;
    int result_finite = 1;
    for (int64_t i = 0; i < C->nnz; ++i)
        if (!isfinite(C->val[i])) result_finite = 0;
    int passed = ref_finite && result_finite;
    for (int32_t r = 0; r < ref_rows; ++r)
        if (ref_nnz[r] != C->row_ptr[r + 1] - C->row_ptr[r]) ++*mismatches;
    return passed && *mismatches == 0;
}

/* ------------------------------------------------------------------------ */
/* Full validation (componentwise_dot_bound_v1): spgemm_common.cpp:590-851.   */
/* ------------------------------------------------------------------------ */
typedef long double ld;

static ld upward(ld v) { return nextafterl(v, (ld)INFINITY); }            /* :596-598 */
static ld add_up(ld a, ld b) { if (a == 0) return b; if (b == 0) return a; return upward(a + b); } /* :599-603 */
static ld mul_up(ld a, ld b) { return a == 0 || b == 0 ? 0 : upward(a * b); } /* :604-606 */
static ld div_up(ld a, ld b) { return a == 0 ? 0 : upward(a / b); }       /* :607-609 */

typedef struct { ld gamma, underflow; } err_model;

/* errorModel<T>: spgemm_common.cpp:652-661; eps2 = epsilon<T>/2, tiny = denorm_min<T> */
static err_model error_model(int64_t terms, ld eps2, ld tiny) {
    const ld operations = mul_up(2, (ld)terms);
    const ld scaled = mul_up(operations, eps2);
    err_model m;
    if (!(scaled < 0.5L)) {
        m.gamma = 1;
        m.underflow = (ld)INFINITY;
        return m;
    }
    const ld denominator = nextafterl(1 - scaled, 0.0L);
    m.gamma = div_up(scaled, denominator);
    m.underflow = div_up(mul_up(operations, tiny), denominator);
    return m;
}

/* coordinateBound: spgemm_common.cpp:663-675 */
static ld coordinate_bound(int64_t terms, ld reference, ld magnitude) {
    if (terms == 0) return 0;
    const err_model cand = error_model(terms, (ld)DBL_EPSILON / 2.0L, (ld)DBL_TRUE_MIN);
    const err_model refm = error_model(terms, LDBL_EPSILON / 2.0L, LDBL_TRUE_MIN);
    if (!(cand.gamma < 1 && refm.gamma < 1) || !isfinite(reference) || !isfinite(magnitude))
        return (ld)INFINITY;
    const ld scale = div_up(add_up(magnitude, refm.underflow), nextafterl(1 - refm.gamma, 0.0L));
    return add_up(mul_up(add_up(cand.gamma, refm.gamma), scale), add_up(cand.underflow, refm.underflow));
}

/* gradualUnderflow<T>: spgemm_common.cpp:611-619 */
static int gradual_underflow_d(void) {
    volatile double normal = DBL_MIN, half = 0.5, sub = DBL_TRUE_MIN, large = DBL_MAX;
    return normal * half == DBL_MIN / 2.0 && sub * large == DBL_TRUE_MIN * DBL_MAX;
}
static int gradual_underflow_ld(void) {
    volatile ld normal = LDBL_MIN, half = 0.5L, sub = LDBL_TRUE_MIN, large = LDBL_MAX;
    return normal * half == LDBL_MIN / 2.0L && sub * large == LDBL_TRUE_MIN * LDBL_MAX;
}

/* fullNumericEnvironment: spgemm_common.cpp:676-693 */
int orc_full_environment(void) {
#if defined(__FAST_MATH__) || (defined(__FINITE_MATH_ONLY__) && __FINITE_MATH_ONLY__ > 0)
    return 0;
#endif
    if (FLT_RADIX != 2 || DBL_MANT_DIG != 53 || LDBL_MANT_DIG < DBL_MANT_DIG) return 0;
    if (fegetround() != FE_TONEAREST || !gradual_underflow_d() || !gradual_underflow_ld()) return 0;
    return 1;
}

/* validNumericCsr: spgemm_common.cpp:621-638 */
static int valid_numeric_csr(const orc_csr *m, int ordered) {
    if (m->rows < 0 || m->cols < 0 || !m->row_ptr) return 0;
    if (m->row_ptr[0] != 0 || m->row_ptr[m->rows] != m->nnz) return 0;
    for (int32_t r = 0; r < m->rows; ++r) {
        const int64_t first = m->row_ptr[r], end = m->row_ptr[r + 1];
        if (first < 0 || end < first || end > m->nnz) return 0;
        int32_t previous = -1;
        for (int64_t p = first; p < end; ++p) {
            const int32_t c = m->col[p];
            if (c < 0 || c >= m->cols || (ordered && c <= previous) || !isfinite(m->val[p])) return 0;
            previous = c;
        }
    }
    return 1;
}

typedef struct {
    int32_t passed;
    int32_t environment_ok;
    int32_t matrices_ok;
    int32_t worst_row;
    int32_t worst_column;
    int32_t reference_digits;
    int64_t rejected;
    int64_t row_nnz_mismatches;
    int64_t support_mismatches;
    int64_t reference_nnz;
    double max_abs_error;        /* max |actual - serial|  (historical TSV field) */
    double max_reference_error;  /* max |actual - long-double reference| */
    double max_error_ratio;      /* max upward(err)/bound, as in the reference */
    double worst_actual, worst_reference, worst_bound;
    /* Added reports (not acceptance criteria): errors against the fp64 serial
     * oracle product, max-entry relative  max|c-s| / max|s|  and Frobenius
     * relative  ||c-s||_F / ||s||_F (0 when both are zero). */
    double max_entry_rel_error;
    double frobenius_rel_error;
} orc_full_report;

typedef struct { int32_t column; double serial; ld reference, bound; int64_t terms; } full_cell;

static int cmp_cell(const void *x, const void *y) {
    const full_cell *a = x, *b = y;
    return (a->column > b->column) - (a->column < b->column);
}

/* validateFullStreaming: spgemm_common.cpp:720-851 (uncached branch; the
 * per-row reference cells are buildFullRow, spgemm_common.cpp:695-717). */
int orc_full_validate(const orc_csr *A, const orc_csr *B, const orc_csr *C, orc_full_report *rep) {
    memset(rep, 0, sizeof *rep);
    rep->max_abs_error = INFINITY;
    rep->reference_digits = LDBL_MANT_DIG;
    rep->worst_row = -1;
    rep->worst_column = -1;
    rep->environment_ok = orc_full_environment();
    if (!rep->environment_ok) return ORC_OK;
    if (!valid_numeric_csr(A, 0) || !valid_numeric_csr(B, 0) || !valid_numeric_csr(C, 1) ||
        A->cols != B->rows || C->rows != A->rows || C->cols != B->cols)
        return ORC_OK;
    rep->matrices_ok = 1;
    size_t ncols = B->cols > 0 ? (size_t)B->cols : 1;
    double *serial = malloc(ncols * sizeof(double));
    ld *refv = malloc(ncols * sizeof(ld));
    ld *mag = malloc(ncols * sizeof(ld));
    int64_t *terms = malloc(ncols * sizeof(int64_t));
    int32_t *mark = malloc(ncols * sizeof(int32_t));
    full_cell *cells = malloc(ncols * sizeof(full_cell));
    if (!serial || !refv || !mag || !terms || !mark || !cells) {
        free(serial); free(refv); free(mag); free(terms); free(mark); free(cells);
        return fail(ORC_ENOMEM, "out of memory%s", "");
    }
    for (size_t j = 0; j < ncols; ++j) mark[j] = -1;
    double max_serial_err = 0;
    ld max_ref_err = 0, max_ratio = 0;
    double diff_max = 0, ref_max = 0;
    ld diff_ss = 0, ref_ss = 0;
    int64_t support = 0, refnnz = 0, rejected = 0, rowmis = 0;
    for (int32_t row = 0; row < A->rows; ++row) {
        /* buildFullRow: spgemm_common.cpp:695-717 */
        int64_t nc = 0;
        for (int64_t a = A->row_ptr[row]; a < A->row_ptr[row + 1]; ++a) {
            const int32_t k = A->col[a];
            for (int64_t b = B->row_ptr[k]; b < B->row_ptr[k + 1]; ++b) {
                const int32_t j = B->col[b];
                if (mark[j] != row) {
                    mark[j] = row;
                    serial[j] = 0;
                    refv[j] = 0;
                    mag[j] = 0;
                    terms[j] = 0;
                    cells[nc++].column = j;
                }
                const double product = A->val[a] * B->val[b];
                const ld precise = (ld)A->val[a] * (ld)B->val[b];
                serial[j] += product;
                refv[j] += precise;
                mag[j] += fabsl(precise);
                terms[j] += 1;
            }
        }
        for (int64_t i = 0; i < nc; ++i) {
            const int32_t j = cells[i].column;
            cells[i].serial = serial[j];
            cells[i].reference = refv[j];
            cells[i].bound = coordinate_bound(terms[j], refv[j], mag[j]);
            cells[i].terms = terms[j];
        }
        qsort(cells, (size_t)nc, sizeof(full_cell), cmp_cell);
        int64_t row_nnz = 0;
        for (int64_t i = 0; i < nc; ++i)
            if (cells[i].serial != 0) ++row_nnz;
        refnnz += row_nnz;
        int64_t actual = C->row_ptr[row];
        const int64_t actual_end = C->row_ptr[row + 1];
        if (actual_end - actual != row_nnz) ++rowmis;
        int64_t expected = 0;
        while (actual < actual_end || expected < nc) {
            full_cell cell = {0, 0, 0, 0, 0};
            double actual_value = 0;
            int stored = 0;
            int32_t column;
            if (expected == nc || (actual < actual_end && C->col[actual] < cells[expected].column)) {
                column = C->col[actual];
                actual_value = C->val[actual++];
                stored = 1;
            } else {
                cell = cells[expected++];
                column = cell.column;
                if (actual < actual_end && C->col[actual] == column) {
                    actual_value = C->val[actual++];
                    stored = 1;
                }
            }
            const int support_mismatch = stored != (cell.serial != 0);
            double serial_err = fabs(actual_value - cell.serial);
            if (!isfinite(serial_err)) serial_err = INFINITY;
            if (serial_err > max_serial_err) max_serial_err = serial_err;
            const ld difference = fabsl((ld)actual_value - cell.reference);
            const ld error_upper = difference == 0 ? 0 : upward(difference);
            const ld bound = cell.bound;
            const int accepted = isfinite(serial_err) && isfinite(error_upper) && isfinite(bound) &&
                                 error_upper <= bound;
            ld ratio;
            if (!isfinite(serial_err) || !isfinite(error_upper) || !isfinite(bound) ||
                (bound == 0 && difference != 0))
                ratio = (ld)INFINITY;
            else if (bound == 0)
                ratio = 0;
            else if (accepted) {
                ratio = div_up(error_upper, bound);
                if (ratio > 1.0L) ratio = 1.0L;
            } else
                ratio = div_up(error_upper, bound);
            if (ratio > max_ratio) {
                max_ratio = ratio;
                rep->worst_row = row;
                rep->worst_column = column;
                rep->worst_actual = actual_value;
                rep->worst_reference = (double)cell.reference;
                rep->worst_bound = (double)bound;
            }
            const ld dfin = isfinite(difference) ? difference : (ld)INFINITY;
            if (dfin > max_ref_err) max_ref_err = dfin;
            if (!accepted) ++rejected;
            if (support_mismatch) ++support;
            /* added reports against the serial fp64 oracle */
            const double d = fabs(actual_value - cell.serial);
            if (d > diff_max || isnan(d)) diff_max = isnan(d) ? INFINITY : d;
            if (fabs(cell.serial) > ref_max) ref_max = fabs(cell.serial);
            diff_ss += (ld)d * (ld)d;
            ref_ss += (ld)cell.serial * (ld)cell.serial;
        }
    }
    free(serial); free(refv); free(mag); free(terms); free(mark); free(cells);
    rep->max_abs_error = max_serial_err;
    rep->max_reference_error = (double)max_ref_err;
    rep->max_error_ratio = (double)max_ratio;
    rep->rejected = rejected;
    rep->row_nnz_mismatches = rowmis;
    rep->support_mismatches = support;
    rep->reference_nnz = refnnz;
    rep->passed = rejected == 0;
    rep->max_entry_rel_error = diff_max == 0 ? 0 : (ref_max > 0 ? diff_max / ref_max : INFINITY);
    rep->frobenius_rel_error = diff_ss == 0 ? 0 : (ref_ss > 0 ? (double)sqrtl(diff_ss / ref_ss) : INFINITY);
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* percentile90: spgemm_common.cpp:916-924 (nearest rank, ceil(0.9 n) - 1)   */
/* ------------------------------------------------------------------------ */
static int cmp_f64(const void *x, const void *y) {
    double a = *(const double *)x, b = *(const double *)y;
    return (a > b) - (a < b);
}

double orc_percentile90(const double *samples, int64_t n) {
    if (n <= 0) return NAN;
    double *s = malloc((size_t)n * sizeof(double));
    if (!s) return NAN;
    memcpy(s, samples, (size_t)n * sizeof(double));
    qsort(s, (size_t)n, sizeof(double), cmp_f64);
    int64_t index = (int64_t)ceil(0.90 * (double)n);
    if (index < 1) index = 1;
    double v = s[index - 1];
    free(s);
    return v;
}

/* ------------------------------------------------------------------------ */
/* trident::LocalAccumulator: matrix_blocks.cpp:152-200.                      */
/*   addProduct: accumulator[col] += a*b, value-initialised to 0.0, for A     */
/*   entries then B entries in stored order, stage after stage;               */
/*   finish: drop entries == 0.0, ascending columns.                          */
/* The per-row maps are held as a CSR with every touched column (zeros        */
/* included) so the next stage continues from exactly the stored values.      */
/* ------------------------------------------------------------------------ */
typedef struct {
    int32_t rows, cols;
    orc_csr cur; /* all touched coordinates, sorted, zeros kept */
} orc_acc;

orc_acc *orc_acc_create(int32_t rows, int32_t cols) {
    orc_acc *a = calloc(1, sizeof *a);
    if (!a) return NULL;
    a->rows = rows;
    a->cols = cols;
    a->cur.rows = rows;
    a->cur.cols = cols;
    a->cur.row_ptr = calloc((size_t)rows + 1, sizeof(int64_t));
    a->cur.col = malloc(sizeof(int32_t));
    a->cur.val = malloc(sizeof(double));
    return a;
}

void orc_acc_destroy(orc_acc *a) {
    if (!a) return;
    orc_free(&a->cur);
    free(a);
}

int orc_acc_add(orc_acc *acc, const orc_csr *A, const orc_csr *B) {
    /* matrix_blocks.cpp:154-157 shape check */
    if (A->cols != B->rows || A->rows != acc->rows || B->cols != acc->cols)
        return fail(ORC_EINVAL, "incompatible Trident partial product shapes%s", "");
    size_t ncols = acc->cols > 0 ? (size_t)acc->cols : 1;
    double *v = malloc(ncols * sizeof(double));
    int32_t *mark = malloc(ncols * sizeof(int32_t));
    int32_t *touched = malloc(ncols * sizeof(int32_t));
    int64_t *rp = calloc((size_t)acc->rows + 1, sizeof(int64_t));
    obuf out = {0, 0, NULL, NULL};
    int status = (v && mark && touched && rp) ? ORC_OK : ORC_ENOMEM;
    for (size_t j = 0; status == ORC_OK && j < ncols; ++j) mark[j] = -1;
    for (int32_t r = 0; status == ORC_OK && r < acc->rows; ++r) {
        int64_t nt_ = 0;
        for (int64_t p = acc->cur.row_ptr[r]; p < acc->cur.row_ptr[r + 1]; ++p) {
            const int32_t j = acc->cur.col[p];
            mark[j] = r;
            v[j] = acc->cur.val[p];
            touched[nt_++] = j;
        }
        for (int64_t e = A->row_ptr[r]; e < A->row_ptr[r + 1]; ++e) {
            const int32_t k = A->col[e];
            if (k < 0 || k >= B->rows) { status = ORC_ERANGE; break; }
            for (int64_t t = B->row_ptr[k]; t < B->row_ptr[k + 1]; ++t) {
                const int32_t j = B->col[t];
                if (mark[j] != r) {
                    mark[j] = r;
                    v[j] = 0.0;
                    touched[nt_++] = j;
                }
                v[j] += A->val[e] * B->val[t];
            }
        }
        if (status != ORC_OK) break;
        qsort(touched, (size_t)nt_, sizeof(int32_t), cmp_i32);
        for (int64_t i = 0; i < nt_; ++i)
            if (!obuf_push(&out, touched[i], v[touched[i]])) { status = ORC_ENOMEM; break; }
        rp[r + 1] = out.n;
    }
    free(v); free(mark); free(touched);
    if (status != ORC_OK) {
        free(rp); free(out.col); free(out.val);
        return fail(status, "accumulate failed%s", "");
    }
    orc_free(&acc->cur);
    acc->cur.row_ptr = rp;
    acc->cur.col = out.col ? out.col : malloc(sizeof(int32_t));
    acc->cur.val = out.val ? out.val : malloc(sizeof(double));
    acc->cur.nnz = out.n;
    return ORC_OK;
}

int orc_acc_finish(const orc_acc *acc, orc_csr *C) {
    memset(C, 0, sizeof *C);
    C->rows = acc->rows;
    C->cols = acc->cols;
    const orc_csr *s = &acc->cur;
    C->row_ptr = calloc((size_t)acc->rows + 1, sizeof(int64_t));
    C->col = malloc((size_t)(s->nnz ? s->nnz : 1) * sizeof(int32_t));
    C->val = malloc((size_t)(s->nnz ? s->nnz : 1) * sizeof(double));
    if (!C->row_ptr || !C->col || !C->val) {
        orc_free(C);
        return fail(ORC_ENOMEM, "out of memory%s", "");
    }
    int64_t n = 0;
    for (int32_t r = 0; r < acc->rows; ++r) {
        for (int64_t p = s->row_ptr[r]; p < s->row_ptr[r + 1]; ++p) {
            if (s->val[p] != 0.0) {
                C->col[n] = s->col[p];
                C->val[n] = s->val[p];
                ++n;
            }
        }
        C->row_ptr[r + 1] = n;
    }
    C->nnz = n;
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* readMatrixMarket: matrix_market.cpp:28-144.                               */
/*   coordinate real/integer/pattern; general/symmetric/skew-symmetric/       */
/*   hermitian (square only); mirror entry emitted right after each           */
/*   off-diagonal entry (skew: negated); 'D'/'d' exponents; file order,       */
/*   duplicates and explicit zeros preserved (stable counting sort by row).   */
/* Token parsing follows the istream rules the reference relies on: tokens    */
/* separated by whitespace, extra tokens rejected, non-finite values          */
/* rejected.                                                                  */
/* ------------------------------------------------------------------------ */
typedef struct { int32_t row, col; double val; } mm_entry;

static int comment_or_empty(const char *line) {
    /* matrix_market.cpp:21-24 */
    const char *p = line;
    while (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n') ++p;
    return *p == '\0' || *p == '%';
}

static char *next_token(char **cursor) {
    char *p = *cursor;
    while (*p && isspace((unsigned char)*p)) ++p;
    if (!*p) { *cursor = p; return NULL; }
    char *start = p;
    while (*p && !isspace((unsigned char)*p)) ++p;
    if (*p) *p++ = '\0';
    *cursor = p;
    return start;
}

static int parse_i64(const char *tok, int64_t *out) {
    if (!tok) return 0;
    errno = 0;
    char *end;
    long long v = strtoll(tok, &end, 10);
    if (end == tok || *end != '\0' || errno == ERANGE) return 0;
    *out = v;
    return 1;
}

static int parse_i32(const char *tok, int32_t *out) {
    int64_t v;
    if (!parse_i64(tok, &v) || v < INT32_MIN || v > INT32_MAX) return 0;
    *out = (int32_t)v;
    return 1;
}

static int parse_f64(const char *tok, double *out) {
    if (!tok) return 0;
    /* istream >> double accepts decimal floating forms only */
    for (const char *p = tok; *p; ++p)
        if (!(isdigit((unsigned char)*p) || *p == '+' || *p == '-' || *p == '.' || *p == 'e' || *p == 'E'))
            return 0;
    char *end;
    double v = strtod(tok, &end);
    if (end == tok || *end != '\0') return 0;
    *out = v;
    return 1;
}

static void lower(char *s) {
    for (; *s; ++s) *s = (char)tolower((unsigned char)*s);
}

static ssize_t read_line(FILE *f, char **buf, size_t *cap) { return getline(buf, cap, f); }

int orc_mm_read(const char *path, orc_csr *out) {
    memset(out, 0, sizeof *out);
    FILE *f = fopen(path, "r");
    if (!f) return fail(ORC_EPARSE, "cannot open Matrix Market file: %s", path);
    char *line = NULL;
    size_t cap = 0;
    mm_entry *ents = NULL;
    int status = ORC_OK;
    if (read_line(f, &line, &cap) < 0) { status = fail(ORC_EPARSE, "Matrix Market file is empty: %s", path); goto done; }
    {
        char *cur = line;
        char *banner = next_token(&cur), *object = next_token(&cur), *storage = next_token(&cur);
        char *field = next_token(&cur), *symmetry = next_token(&cur), *extra = next_token(&cur);
        char e[1] = {0};
        if (!object) object = e;
        if (!storage) storage = e;
        if (!field) field = e;
        if (!symmetry) symmetry = e;
        lower(object); lower(storage); lower(field); lower(symmetry);
        if (!banner || strcmp(banner, "%%MatrixMarket") != 0 || strcmp(object, "matrix") != 0 ||
            strcmp(storage, "coordinate") != 0 || extra) {
            status = fail(ORC_EPARSE, "only Matrix Market coordinate matrices are supported: %s", path);
            goto done;
        }
        int pattern = strcmp(field, "pattern") == 0;
        if (strcmp(field, "real") != 0 && strcmp(field, "integer") != 0 && !pattern) {
            status = fail(ORC_EPARSE, "only real, integer, and pattern Matrix Market fields are supported: %s", path);
            goto done;
        }
        int skew = strcmp(symmetry, "skew-symmetric") == 0;
        int general = strcmp(symmetry, "general") == 0;
        if (!general && strcmp(symmetry, "symmetric") != 0 && !skew && strcmp(symmetry, "hermitian") != 0) {
            status = fail(ORC_EPARSE, "unsupported Matrix Market symmetry: %s", symmetry);
            goto done;
        }
        do {
            if (read_line(f, &line, &cap) < 0) { status = fail(ORC_EPARSE, "Matrix Market file has no size line: %s", path); goto done; }
        } while (comment_or_empty(line));
        int32_t rows, cols;
        int64_t entries;
        cur = line;
        char *t1 = next_token(&cur), *t2 = next_token(&cur), *t3 = next_token(&cur), *t4 = next_token(&cur);
        if (!parse_i32(t1, &rows) || !parse_i32(t2, &cols) || !parse_i64(t3, &entries) || rows <= 0 ||
            cols <= 0 || entries < 0 || t4) {
            status = fail(ORC_EPARSE, "invalid Matrix Market dimensions: %s", path);
            goto done;
        }
        const int mirrored = !general;
        if (mirrored && rows != cols) {
            snprintf(g_err, sizeof g_err, "Matrix Market %s storage requires a square matrix: %s", symmetry, path);
            status = ORC_EPARSE;
            goto done;
        }
        int64_t capn = mirrored ? 2 * entries : entries;
        ents = malloc((size_t)(capn ? capn : 1) * sizeof(mm_entry));
        if (!ents) { status = fail(ORC_ENOMEM, "out of memory%s", ""); goto done; }
        int64_t n = 0;
        for (int64_t i = 0; i < entries; ++i) {
            do {
                if (read_line(f, &line, &cap) < 0) { status = fail(ORC_EPARSE, "unexpected end of Matrix Market entries: %s", path); goto done; }
            } while (comment_or_empty(line));
            for (char *p = line; *p; ++p) {
                if (*p == 'D') *p = 'E';
                else if (*p == 'd') *p = 'e';
            }
            cur = line;
            int32_t r, c;
            double v = 1.0;
            char *a1 = next_token(&cur), *a2 = next_token(&cur);
            int ok = parse_i32(a1, &r) && parse_i32(a2, &c);
            if (ok && !pattern) ok = parse_f64(next_token(&cur), &v);
            if (!ok || !isfinite(v) || next_token(&cur)) { status = fail(ORC_EPARSE, "invalid Matrix Market entry: %s", path); goto done; }
            if (r < 1 || r > rows || c < 1 || c > cols) { status = fail(ORC_EPARSE, "Matrix Market entry index is out of range: %s", path); goto done; }
            --r;
            --c;
            if (skew && r == c && v != 0.0) { status = fail(ORC_EPARSE, "skew-symmetric diagonal must be zero: %s", path); goto done; }
            ents[n].row = r; ents[n].col = c; ents[n].val = v; ++n;
            if (mirrored && r != c) {
                ents[n].row = c; ents[n].col = r; ents[n].val = skew ? -v : v; ++n;
            }
        }
        while (read_line(f, &line, &cap) >= 0) {
            if (!comment_or_empty(line)) { status = fail(ORC_EPARSE, "more Matrix Market entries than declared: %s", path); goto done; }
        }
        out->rows = rows;
        out->cols = cols;
        out->nnz = n;
        out->row_ptr = calloc((size_t)rows + 1, sizeof(int64_t));
        out->col = malloc((size_t)(n ? n : 1) * sizeof(int32_t));
        out->val = malloc((size_t)(n ? n : 1) * sizeof(double));
        int64_t *cursor = malloc(((size_t)rows + 1) * sizeof(int64_t));
        if (!out->row_ptr || !out->col || !out->val || !cursor) {
            free(cursor);
            orc_free(out);
            status = fail(ORC_ENOMEM, "out of memory%s", "");
            goto done;
        }
        for (int64_t i = 0; i < n; ++i) out->row_ptr[ents[i].row + 1]++;
        for (int32_t r = 0; r < rows; ++r) out->row_ptr[r + 1] += out->row_ptr[r];
        memcpy(cursor, out->row_ptr, ((size_t)rows + 1) * sizeof(int64_t));
        for (int64_t i = 0; i < n; ++i) {
            const int64_t pos = cursor[ents[i].row]++;
            out->col[pos] = ents[i].col;
            out->val[pos] = ents[i].val;
        }
        free(cursor);
    }
done:
    free(line);
    free(ents);
    fclose(f);
    return status;
}
