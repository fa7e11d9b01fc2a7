/*
 * spgemm_amd.h -- C ABI of the MI355X SpGEMM engine (libspgemm_amd.so).
 *
 * The engine replaces the local multiply each rank of malchioman/SpGEMM-hpc-thesis
 * computes: A_local rows times the gathered B rows.  The data contract is the
 * reference's CsrMatrix (src/common/csr_matrix.hpp:10-49): int64 row offsets and
 * entry counts, int32 column indices, fp64 values.
 *
 * Results are bit-identical to the reference's serial order: every output
 * coordinate is the left-to-right sum of a*b over A entries then B entries in
 * stored order, columns ascending, exact zeros dropped.
 *
 * All entry points return an spg_status; spg_last_error() gives the message of
 * the last failure on a context.  One context per host thread; calls on one
 * context are synchronous with respect to the host unless noted.
 */
#ifndef SPGEMM_AMD_H
#define SPGEMM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes.  The C++ shim (spgemm-hpc-thesis_amd/host/spgemm_local.hpp)
 * maps them back to the reference's exception types:
 *   SPG_EINVAL       -> std::invalid_argument ("SpGEMM requires A.cols == B.rows",
 *                       src/common/spgemm_common.cpp:106-110; malformed CSR views)
 *   SPG_EOVERFLOW    -> std::overflow_error (src/common/csr_matrix.hpp:13-41)
 *   SPG_ERANGE       -> std::runtime_error ("A contains a column index outside B's
 *                       row range", src/common/spgemm_common.cpp:141-143): batched
 *                       and served products, whose semantics are the serial
 *                       product's
 *   SPG_EMISSING_ROW -> fail("missing a remote sparse row required by A"),
 *                       src/baselines/spgemm_exchange.cpp:82-85: the drop-in
 *                       multiply, for every A column outside the local block
 *                       and the halo
 *   SPG_ENOMEM, SPG_EDEVICE, SPG_ENODEVICE, SPG_ESTATE, SPG_EPARSE ->
 *                       std::runtime_error / std::bad_alloc */
typedef enum spg_status {
    SPG_OK = 0,
    SPG_EINVAL = 1,
    SPG_EOVERFLOW = 2,
    SPG_ENOMEM = 3,
    SPG_EMISSING_ROW = 4,
    SPG_ERANGE = 5,
    SPG_EDEVICE = 6,
    SPG_ENODEVICE = 7,
    SPG_ESTATE = 8,
    SPG_EPARSE = 9,
    SPG_EBUSY = 10
} spg_status;

typedef enum spg_location { SPG_HOST = 0, SPG_DEVICE = 1 } spg_location;

/* A borrowed CSR matrix (the reference's CsrMatrix, csr_matrix.hpp:43-49).
 * row_ptr has rows+1 entries with row_ptr[0] == 0 and row_ptr[rows] == nnz.
 * Pointers are host or device memory according to `location`. */
typedef struct spg_csr_view {
    int32_t rows;
    int32_t cols;
    int64_t nnz;
    const int64_t *row_ptr;
    const int32_t *col_idx;
    const double *values;
    int32_t location;
} spg_csr_view;

/* The gathered remote B rows (RemoteSparseRows, src/baselines/spgemm_exchange.hpp:8-20):
 * slot s holds global B row global_rows[s] with entries
 * [row_ptr[s], row_ptr[s+1]).  global_rows must be distinct. */
typedef struct spg_halo_view {
    int32_t n_rows;
    const int32_t *global_rows;
    const int64_t *row_ptr;
    const int32_t *col_idx;
    const double *values;
    int32_t location;
} spg_halo_view;

typedef struct spg_context spg_context;

/* Per-multiply statistics of the last completed product on a context. */
typedef struct spg_stats {
    int64_t products;          /* P = sum over A entries of nnz(B row): the flop basis /2
                                  (scalarMultiplicationCount, spgemm_common.cpp:474-487) */
    int64_t structural_nnz;    /* distinct output coordinates */
    int64_t nnz;               /* stored output entries after exact-zero drop */
    int64_t rows_per_bin[8];   /* rows per kernel bin, by products per row: empty,
                                  <=128, <=256, <=512, <=1024, <=2048, large (chunked) */
    int32_t compacted;         /* 1 if exact zeros forced a compaction pass */
    int32_t bins;              /* number of bins used */
    int32_t single_pass;       /* 1 if the small rows took the single-pass kernel */
    int32_t wide_format;       /* 1 if the 64-bit entry stream and B addressing ran
                                  (B of 2^29 entries or more, or SPG_FORMAT_WIDE) */
    double ms_total;           /* device time of the whole pipeline (events), if timing on */
    double ms_symbolic;        /* device time of the symbolic phase, if timing on */
    double ms_numeric;         /* device time of the numeric phase, if timing on */
    int32_t direct;            /* 1 if every row fit the single pass by the longest rows of A
                                  and B alone, so the per-row analysis was skipped */
    int32_t partitioned;       /* 1 if large rows were partitioned by column window (spg_set_large) */
} spg_stats;

/* ---- context ---------------------------------------------------------- */
spg_status spg_create(spg_context **ctx, int device);
spg_status spg_destroy(spg_context *ctx);
const char *spg_status_string(spg_status status);
const char *spg_last_error(const spg_context *ctx);
/* Kernels of this context run on `stream` (a hipStream_t; NULL = the context's
 * own stream). */
spg_status spg_set_stream(spg_context *ctx, void *stream);
/* The context's current stream (a hipStream_t). */
spg_status spg_context_stream(const spg_context *ctx, void **stream);
/* Make the context's device current on the calling host thread. */
spg_status spg_context_bind(spg_context *ctx);
/* Which pipeline computes the small rows (<= 256 products and A entries):
 *   SPG_PATH_AUTO         single pass when its C capacity (small rows' product
 *                         bounds + big rows' exact counts) fits half of the
 *                         device memory, else two-phase
 *   SPG_PATH_SINGLE_PASS  always single pass
 *   SPG_PATH_TWO_PHASE    count every row first, then compute
 *   SPG_PATH_SINGLE_PASS_DEALT  single pass over 32-row tiles whose small rows
 *                         are dealt to 16 waves: for products whose rows are
 *                         mostly bigger than 256 products
 *   SPG_PATH_SINGLE_PASS_PACKED  single pass over tiles of 16 small rows each
 *                         (per 256-row block, in row order), the other rows'
 *                         counts folded into the tiles: the same use
 * Results are identical on every path.  No counterpart in the reference (its
 * spgemm() has one CPU pipeline); exposed so callers and tests can pin a path. */
enum {
    SPG_PATH_AUTO = 0,
    SPG_PATH_SINGLE_PASS = 1,
    SPG_PATH_TWO_PHASE = 2,
    SPG_PATH_SINGLE_PASS_DEALT = 3,
    SPG_PATH_SINGLE_PASS_PACKED = 4
};
spg_status spg_set_path(spg_context *ctx, int path);
/* Encoding of the per-A-entry B-row stream and of B addressing in the
 * single-pass kernel:
 *   SPG_FORMAT_AUTO  32-bit B-row starts and byte offsets when B (local block
 *                    plus halo) has fewer than 2^29 entries, else 64-bit
 *   SPG_FORMAT_WIDE  always the 64-bit kernels, those that large B operands
 *                    and big batches take, so that small products reach them
 * Results are identical in both formats.  Batches and accumulators created
 * from the context inherit the setting.  No counterpart in the reference. */
enum { SPG_FORMAT_AUTO = 0, SPG_FORMAT_WIDE = 1 };
spg_status spg_set_format(spg_context *ctx, int format);
/* Value kernels of the rows with more than 2048 scalar products:
 *   SPG_LARGE_AUTO     rows whose distinct columns fit an LDS accumulator
 *                      fold there (C with at most 2^16 columns); the others
 *                      are partitioned by windows of 4096 columns into HBM
 *                      scratch and each window folds in LDS, when C has at
 *                      most 2^20 columns and the scratch (10 B per product)
 *                      fits a quarter of the device; else they fold onto
 *                      dense accumulator windows in HBM
 *   SPG_LARGE_WINDOWS  never partitioned (the HBM windows)
 * Results are identical.  Batches and accumulators created from the context
 * inherit the setting; spg_stats.partitioned reports when rows were
 * partitioned.  No counterpart in the reference. */
enum { SPG_LARGE_AUTO = 0, SPG_LARGE_WINDOWS = 1 };
spg_status spg_set_large(spg_context *ctx, int mode);
/* Record HIP events around the phases of each multiply (spg_stats.ms_*). */
spg_status spg_set_timing(spg_context *ctx, int enabled);
spg_status spg_last_stats(const spg_context *ctx, spg_stats *out);

/* ---- the drop-in local multiply ----------------------------------------
 * Replaces
 *   CsrMatrix spgemm(const CsrMatrix& localMatrixA, RowBlock localBBlock,
 *                    const CsrMatrix& localMatrixB, const RemoteSparseRows& remoteBRows,
 *                    MPI_Comm communicator);
 * (src/baselines/spgemm_exchange.hpp:37-39, body spgemm_exchange.cpp:258-325),
 * called at src/baselines/two_sided_spgemm.cpp:108-109,130-131,153-154,
 * one_sided_get_spgemm.cpp:208-209,230-231,252-253 and
 * one_sided_put_spgemm.cpp:327-328,352-354,378-380.
 *
 * B row k of the product is local row k - b_first_row when
 * b_first_row <= k < b_first_row + B_local->rows, else the halo slot holding
 * global row k (bRowForColumn, spgemm_exchange.cpp:75-87).  With halo == NULL
 * and b_first_row == 0 this is the -np 1 product.  As in the reference's
 * spgemm(), the shapes are not compared: an A column that names no local row
 * and no halo slot (a column >= B_local->rows at -np 1, or a negative one) is
 * SPG_EMISSING_ROW, the reference's fail("missing a remote sparse row required
 * by A") (spgemm_exchange.cpp:82-85).  The product is computed on the device;
 * *nnz_c receives nnz(C).  The result stays on the context until the next
 * multiply. */
spg_status spg_multiply_begin(spg_context *ctx, const spg_csr_view *A, const spg_csr_view *B_local,
                              int32_t b_first_row, const spg_halo_view *halo, int64_t *nnz_c);
/* Rows [row_begin, row_end) of the same product: C's rows row_begin.. as a
 * (row_end - row_begin)-row CSR (row_ptr starting at 0), bit-identical to
 * those rows of the whole product.  For outputs larger than device memory,
 * taken in row chunks (the reference computes per-rank row blocks of A,
 * sliceRows in src/common/spgemm_common.cpp; here the chunks belong to one
 * device).  No direct counterpart in the reference's spgemm(). */
spg_status spg_multiply_rows(spg_context *ctx, const spg_csr_view *A, const spg_csr_view *B_local,
                             int32_t b_first_row, const spg_halo_view *halo, int32_t row_begin, int32_t row_end,
                             int64_t *nnz_c);
/* The product of device views A and B (the -np 1 product of
 * spg_multiply_begin: b_first_row 0, no halo) without any device memory
 * allocation.  No counterpart in the reference: for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops).  C goes
 * straight into caller-owned device buffers: row_ptr (A->rows + 1 entries),
 * col_idx and values (`capacity` entries each); *nnz_c receives nnz(C).  The
 * small rows always take the single-pass kernel (with the context's tiling),
 * and results are bit-identical to spg_multiply_begin's.  Every intermediate
 * array lives in a fixed device array of 512 MiB (2^29 bytes) that each GPU
 * the library runs on holds until the process ends; a product whose arrays
 * need more (about 300 MB for 2^20 rows with 2^24 A entries, plus 12 bytes
 * per entry of C's rows of more than 256 products) is SPG_ENOMEM.  So is an
 * output capacity below C's symbolic bound (the small rows' product counts
 * plus the other rows' exact counts), which *nnz_c then receives.  All calls
 * on one GPU share that array: calls that overlap on one GPU race.  The
 * context's last result (spg_multiply_finish) is not set. */
spg_status spg_multiply_noalloc(spg_context *ctx, const spg_csr_view *A, const spg_csr_view *B, int64_t *row_ptr,
                                int32_t *col_idx, double *values, int64_t capacity, int64_t *nnz_c);
/* Copy the result of the last multiply into caller-owned host buffers of
 * sizes A->rows+1, nnz_c and nnz_c (the shim resizes its std::vectors first). */
spg_status spg_multiply_finish(spg_context *ctx, int64_t *row_ptr, int32_t *col_idx, double *values);
/* Rows [row_begin, row_end) of the last result into host buffers: row_ptr
 * (row_end - row_begin + 1 entries, starting at 0) and up to `capacity`
 * entries; *nnz_rows receives the rows' entry count (SPG_EINVAL with the
 * count when capacity is too small).  For spot checks and streaming of large
 * outputs; no counterpart in the reference. */
spg_status spg_result_rows(spg_context *ctx, int32_t row_begin, int32_t row_end, int64_t *row_ptr, int32_t *col_idx,
                           double *values, int64_t capacity, int64_t *nnz_rows);
/* Device view of the last result (pointers owned by the context, valid until
 * the next multiply on it or spg_destroy). */
spg_status spg_result_device(spg_context *ctx, spg_csr_view *out);

/* ---- HBM-resident operands ----------------------------------------------
 * Upload a host CSR matrix once and keep it in device memory owned by the
 * library (for repeated, batched or served products).  spg_device_matrix_view
 * gives a device view (location SPG_DEVICE) usable as A or B_local in
 * spg_multiply_begin.  The handle must be freed before its context. */
typedef struct spg_device_matrix spg_device_matrix;
spg_status spg_matrix_upload(spg_context *ctx, const spg_csr_view *host, spg_device_matrix **out);
spg_status spg_device_matrix_view(const spg_device_matrix *m, spg_csr_view *out);
spg_status spg_device_matrix_free(spg_device_matrix *m);
/* Wait for all work queued on the context's stream. */
spg_status spg_synchronize(spg_context *ctx);

/* ---- batched products ----------------------------------------------------
 * No counterpart in the reference (each rank multiplies once); this is the
 * batched mode of the engine.  A batch packs n independent products
 * C_i = A_i * B_i into one resident arena (the B_i stacked, the A_i stacked
 * with their columns shifted to their B_i's rows) and computes all of them
 * with one device pipeline.  Each C_i is bit-identical to its standalone
 * product.  The batch owns its result buffers and runs on the creating
 * context's stream.  Operands are all host views (packed on the host and
 * uploaded once) or all device views (packed on the device; they need to stay
 * valid only during the call). */
typedef struct spg_batch spg_batch;
spg_status spg_batch_create(spg_context *ctx, int32_t n, const spg_csr_view *A, const spg_csr_view *B,
                            spg_batch **out);
/* Replace a batch's products (same rules as spg_batch_create), reusing its
 * device buffers; its previous results are discarded. */
spg_status spg_batch_rebuild(spg_batch *b, int32_t n, const spg_csr_view *A, const spg_csr_view *B);
/* Chained batch: A_i is the result of product src_index[i] of batch `src`
 * (which must have run); the second product of (R.A).P.  spg_batch_refresh
 * copies the source's current results into this batch's A operands on the
 * device (the source's result structure must be unchanged). */
spg_status spg_batch_create_chained(spg_context *ctx, const spg_batch *src, int32_t n, const int32_t *src_index,
                                    const spg_csr_view *B, spg_batch **out);
spg_status spg_batch_refresh(spg_batch *b);
/* Run every product; *nnz_total receives the entries of all C_i. */
spg_status spg_batch_run(spg_batch *b, int64_t *nnz_total);
/* Run only the count phase (analyze and symbolic kernels) of the next run:
 * *value_bytes receives the device bytes its value phase will add to the
 * batch's buffers (C's entries at 12 B each, exact for rows the symbolic
 * kernels count, the work bound for the single-pass rows, plus the large
 * rows' accumulator windows and the single-pass tile state, less what the
 * buffers already hold from earlier runs).  The next spg_batch_run continues
 * from these counts (eagerly, without a captured graph).  For admission by
 * C's size before it is allocated. */
spg_status spg_batch_count(spg_batch *b, uint64_t *value_bytes);
/* After a first run, replay the planned pipeline from a captured HIP graph
 * (no host synchronisation inside the pipeline). */
spg_status spg_batch_set_graph(spg_batch *b, int enabled);
/* Time the pipeline variants over `repeats` runs each and keep the fastest;
 * *choice receives its index: 0 single pass with 2 rows per wave, 1 with 1
 * row per wave, 2 two-phase, 3 single pass over 32-row tiles whose
 * single-pass rows are dealt to 16 waves, 4 single pass over packed tiles of
 * 16 single-pass rows. */
spg_status spg_batch_autotune(spg_batch *b, int repeats, int *choice);
/* Results of the last run: entry count of product i, and a host copy of C_i
 * (row_ptr: rows_i + 1 entries starting at 0). */
spg_status spg_batch_product_nnz(const spg_batch *b, int32_t i, int64_t *nnz);
spg_status spg_batch_copy(spg_batch *b, int32_t i, int64_t *row_ptr, int32_t *col_idx, double *values);
spg_status spg_batch_info(const spg_batch *b, int32_t i, int32_t *rows, int32_t *cols);
spg_status spg_batch_last_stats(const spg_batch *b, spg_stats *out);
spg_status spg_batch_destroy(spg_batch *b);
/* Run the batch on `stream` (NULL: the creating context's stream). */
spg_status spg_batch_set_stream(spg_batch *b, void *stream);
/* Device bytes the batch holds: its arena, workspace and result buffers
 * (buffers grow and are kept across runs and rebuilds). */
spg_status spg_batch_device_bytes(const spg_batch *b, uint64_t *bytes);
/* Free every device buffer of the batch; it holds no products afterwards
 * and is reused with spg_batch_rebuild (not for chained batches). */
spg_status spg_batch_trim(spg_batch *b);
/* Device bytes that building products of these total sizes into `reuse`
 * (NULL: a new batch of ctx) and running them with C's capacity at c_entries
 * would add to the batch's buffers: its arena (rows and a_nnz of the stacked
 * A_i, b_rows and b_nnz of the stacked B_i), its count workspace (cols: the
 * largest C column count) and C (none when c_entries is 0: the count phase
 * alone).  The large rows' accumulator windows, the
 * single-pass tile state and compaction buffers are not included. */
spg_status spg_batch_estimate(spg_context *ctx, const spg_batch *reuse, int64_t rows, int64_t a_nnz, int64_t b_rows,
                              int64_t b_nnz, int32_t cols, int64_t c_entries, uint64_t *bytes);
/* Message of the batch's last failure (the batch computes on its own
 * context, so its errors are not reported on the creating one). */
const char *spg_batch_last_error(const spg_batch *b);

/* ---- served products ------------------------------------------------------
 * No counterpart in the reference; the serving mode of the engine.  Requests
 * (A, B) are submitted from any host thread and copied; a builder thread
 * admits them in arrival order while the bytes known at submission (operand
 * copies and the count workspace) fit budget_bytes, packs them into batch
 * arenas, uploads them and runs the count phase (analyze and symbolic kernels)
 * on its own stream, then charges C's bytes as the counts give them before C
 * is allocated, waiting for releases until they fit (a batch that cannot fit
 * is split; a single product that cannot fails with SPG_ENOMEM).  A runner
 * thread computes the values on a second stream (continuous batching: the
 * next batch is admitted, uploaded and counted while the current one
 * computes).  Results stay in device memory, charged to the budget, until
 * released; each is bit-identical to computing the product alone. */
typedef struct spg_server spg_server;
spg_status spg_server_create(int device, uint64_t budget_bytes, int max_batch, spg_server **out);
/* Host views (copied); SPG_ENOMEM when the request's operands and count
 * workspace alone exceed the budget, SPG_ERANGE when an A column lies outside
 * B's rows.  spg_server_wait reports SPG_ENOMEM for a request whose counted
 * result cannot fit the budget. */
spg_status spg_server_submit(spg_server *s, const spg_csr_view *A, const spg_csr_view *B, int64_t *request_id);
/* Register a matrix once (uploaded, resident, charged to the budget) and
 * submit products of registered matrices by handle. */
spg_status spg_server_register(spg_server *s, const spg_csr_view *host, int64_t *handle);
spg_status spg_server_unregister(spg_server *s, int64_t handle);
spg_status spg_server_submit_resident(spg_server *s, int64_t a_handle, int64_t b_handle, int64_t *request_id);
/* Block until request `id` is computed; *latency_s = completion - submission. */
spg_status spg_server_wait(spg_server *s, int64_t id, int64_t *nnz, double *latency_s);
/* Copy a completed result to host buffers sized from spg_server_wait's nnz
 * (row_ptr: A.rows + 1 entries). */
spg_status spg_server_result(spg_server *s, int64_t id, int64_t *row_ptr, int32_t *col_idx, double *values);
/* Free a completed request's device result and its budget share. */
spg_status spg_server_release(spg_server *s, int64_t id);
spg_status spg_server_stats(spg_server *s, int64_t *batches, int64_t *completed, double *busy_s,
                            uint64_t *charged_bytes);
spg_status spg_server_destroy(spg_server *s);

/* ---- staged accumulation (trident::LocalAccumulator) -------------------
 * Replaces trident::LocalAccumulator{LocalAccumulator(rows, cols);
 * addProduct(a, b); finish();} (src/trident/common/matrix_blocks.hpp:23-32,
 * matrix_blocks.cpp:152-200), called at src/trident/common/execution.cpp:74-89.
 * Each output coordinate is acc += a*b over stages in call order, then A
 * entries, then B entries in stored order; finish() drops exact zeros and
 * sorts columns. */
typedef struct spg_accumulator spg_accumulator;
spg_status spg_accumulator_create(spg_context *ctx, int32_t rows, int32_t cols, spg_accumulator **acc);
spg_status spg_accumulator_add(spg_accumulator *acc, const spg_csr_view *A, const spg_csr_view *B);
spg_status spg_accumulator_finish(spg_accumulator *acc, int64_t *nnz_c);
spg_status spg_accumulator_copy(spg_accumulator *acc, int64_t *row_ptr, int32_t *col_idx, double *values);
spg_status spg_accumulator_destroy(spg_accumulator *acc);

/* ---- host matrices: MatrixMarket and binary CSR -------------------------
 * readMatrixMarket (src/common/matrix_market.cpp:28-144): coordinate
 * real/integer/pattern, general/symmetric/skew-symmetric/hermitian, file order,
 * duplicates and explicit zeros kept.  The matrix is owned by the library
 * until spg_matrix_free. */
typedef struct spg_matrix spg_matrix;
spg_status spg_matrix_read_mm(const char *path, spg_matrix **out);
spg_status spg_matrix_view(const spg_matrix *m, spg_csr_view *out);
spg_status spg_matrix_free(spg_matrix *m);
/* Binary CSR ("SPGCSR01": int32 rows, int32 cols, int64 nnz, row_ptr, col,
 * val; little endian): written from a host view, read back at disk speed.
 * No counterpart in the reference (it re-parses MatrixMarket every run). */
spg_status spg_matrix_write_csr(const spg_csr_view *m, const char *path);
spg_status spg_matrix_read_csr(const char *path, spg_matrix **out);
/* Last error of the host matrix helpers (thread-local). */
const char *spg_matrix_last_error(void);

/* ---- halo rows in HBM -------------------------------------------------------
 * spg_multiply_begin / spg_multiply_rows also take a device view of B_local
 * together with a device halo view (both SPG_DEVICE): the global-row
 * directory (bRowForColumn, src/baselines/spgemm_exchange.cpp:75-87) is then
 * built on the device, so rows received straight into HBM never cross PCIe.
 * A library-owned halo comes from spg_halo_upload (host rows) or
 * spg_halo_pack (the sender's side). */
typedef struct spg_device_halo spg_device_halo;
spg_status spg_halo_upload(spg_context *ctx, const spg_halo_view *host, spg_device_halo **out);
/* The requested global rows of the local block B_local (a device view whose
 * global rows start at first_row), packed in request order as the rows the
 * two-sided exchange sends (prepareRequestedRows + the packing loop of
 * exchangeRemoteSparseRowsTwoSided, spgemm_exchange.cpp:60-73, 217-235): slot s
 * holds row requests[s].  A request outside the local block is
 * SPG_EMISSING_ROW ("received a request for a non-local sparse row of B").
 * requests_location says where `requests` lives (SPG_HOST / SPG_DEVICE). */
spg_status spg_halo_pack(spg_context *ctx, const spg_csr_view *B_local, int32_t first_row, const int32_t *requests,
                         int32_t n, int32_t requests_location, spg_device_halo **out);
spg_status spg_device_halo_view(const spg_device_halo *h, spg_halo_view *out);
spg_status spg_device_halo_free(spg_device_halo *h);

/* ---- validation of a product and its reference files --------------------
 * Host code, multithreaded over rows (threads <= 0: all hardware threads);
 * every decision and diagnostic equals the serial traversal's.  Views must be
 * SPG_HOST.  Errors: spg_validation_last_error (thread-local). */
const char *spg_validation_last_error(void);

/* One reported coordinate (the first eight in row-major order whose support
 * differs or that are rejected), as printed by validation_difference lines. */
typedef struct spg_full_example {
    int32_t row, column;
    int32_t accepted, support_mismatch;
    int64_t contributions;
    double actual, serial_expected;
    long double expected, abs_error, error_bound, error_ratio;
} spg_full_example;

typedef struct spg_full_report {
    int32_t environment_ok; /* binary64 + extended, round to nearest, gradual underflow */
    int32_t invalid_matrix; /* an operand or C failed validNumericCsr, or shapes differ */
    int32_t passed;         /* no rejected coordinate (never set without both checks above) */
    int32_t reference_digits;
    int32_t worst_row, worst_column; /* -1: no coordinate with a positive ratio */
    int32_t n_examples, reserved_;
    int64_t rows, reference_nnz, row_nnz_mismatches, support_mismatches, rejected_coordinates;
    double max_abs_error;
    long double max_reference_error, max_error_ratio, worst_actual, worst_reference, worst_bound;
    spg_full_example examples[8];
} spg_full_report;

/* Full validation, criterion componentwise_dot_bound_v1: replaces
 * validateFullStreaming (src/common/spgemm_common.cpp:720-851).  Expected rows
 * come from the .sfull file `sfull_path` named by `key` (64 hex digits), or,
 * with sfull_path NULL, are rebuilt from A and B (buildFullRow, :694-717). */
spg_status spg_validate_full(const spg_csr_view *A, const spg_csr_view *B, const spg_csr_view *C,
                             const char *sfull_path, const char *key, int threads, spg_full_report *rep);
/* The reference's text lines for a report (validation_* key=value); *needed
 * receives the size including the terminating zero. */
spg_status spg_full_report_text(const spg_full_report *rep, char *buf, size_t cap, size_t *needed);

/* Partial validation, criterion row_nnz_exact_v1: replaces
 * buildPartialReference / comparePartialReference (:541-571).  Counts come from
 * the .spref file `spref_path` (named by `key`) or, with NULL, are rebuilt. */
spg_status spg_partial_counts(const spg_csr_view *A, const spg_csr_view *B, int threads, int64_t *row_nnz,
                              int32_t *finite);
spg_status spg_validate_partial(const spg_csr_view *A, const spg_csr_view *B, const spg_csr_view *C,
                                const char *spref_path, const char *key, int threads, int32_t *passed,
                                int64_t *row_nnz_mismatches);

/* .spref files (encode/decodePartialReference, src/common/partial_reference.cpp:202-245). */
spg_status spg_spref_write(const char *path, const char *key, int32_t rows, int32_t cols, const int64_t *row_nnz,
                           int32_t finite);
spg_status spg_spref_read(const char *path, const char *key, int32_t rows, int32_t cols, int64_t *row_nnz,
                          int32_t *finite);

/* .sfull files (src/common/full_reference.cpp:147-235): written from A and B
 * (prepareFullReference), read row by row (FullReferenceReader); finish checks
 * the SHA-256 trailer and must succeed before a PASS is reported. */
typedef struct spg_sfull spg_sfull;
spg_status spg_sfull_write(const spg_csr_view *A, const spg_csr_view *B, const char *path, const char *key,
                           int threads);
spg_status spg_sfull_open(const char *path, const char *key, int32_t rows, int32_t cols, spg_sfull **out);
spg_status spg_sfull_next_row(spg_sfull *s, int64_t *ncells);
spg_status spg_sfull_row_cells(const spg_sfull *s, int32_t *column, double *serial, long double *reference,
                               long double *bound, int64_t *terms);
spg_status spg_sfull_finish(spg_sfull *s);
void spg_sfull_close(spg_sfull *s);

/* SHA-256 of a byte string as 64 hex digits (the files' checksums). */
spg_status spg_sha256_hex(const void *data, size_t n, char out[65]);

#ifdef __cplusplus
}
#endif

#endif /* SPGEMM_AMD_H */
