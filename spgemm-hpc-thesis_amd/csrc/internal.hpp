// internal.hpp -- shared declarations between the HIP kernels and the host engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace spg {

// Row bins, by the row's upper-bound work P_row (sum of the lengths of the B
// rows its A entries reference) and its A-entry count.
enum Bin : int {
    kBinEmpty = 0,  // P_row == 0: no kernel
    kBinW1 = 1,     // one 64-lane workgroup per row, <= kCapW1 products
    kBinW2 = 2,     // one 64-lane workgroup per row, <= kCapW2 products
    kBinW3 = 3,     // one 64-lane workgroup per row, <= kCapW3 products
    kBinW4 = 4,     // one 128-lane workgroup per row, <= kCapW4 products
    kBinBlock = 5,  // one 256-lane workgroup per row, <= kCapBlock products
    kBinLarge = 6,  // chunked fold onto a windowed dense accumulator in HBM
    kNumBins = 7
};

constexpr int kCapW1 = 128;
constexpr int kCapW2 = 256;
constexpr int kCapW3 = 512;
// Bin W4 holds rows of up to 1024 products in half the LDS and registers per
// row of bin Block, so twice as many rows are resident per CU (six against
// three, for the values and for the counts).
constexpr int kCapW4 = 1024;
constexpr int kTpbW4 = 128;
constexpr int kCapBlock = 2048;
constexpr int kTpbBlock = 256;
// Workgroup sizes of the symbolic kernels for bins W3, W4 and Block: ITEMS =
// capacity / size products per thread.  128 for W3 measured 11 % faster than
// 64 on batched products (more resident waves per CU); wider was slower.
#ifndef SPG_SYM_TPB_W3
#define SPG_SYM_TPB_W3 128
#endif
#ifndef SPG_SYM_TPB_BLOCK
#define SPG_SYM_TPB_BLOCK 256
#endif
#ifndef SPG_SYM_TPB_W4
#define SPG_SYM_TPB_W4 256
#endif
constexpr int kSymTpbW3 = SPG_SYM_TPB_W3, kSymTpbW4 = SPG_SYM_TPB_W4, kSymTpbBlock = SPG_SYM_TPB_BLOCK;
// Large-row path: column windows of up to kLdsBitsCols columns keep the
// presence bitmap in LDS (32 KB), wider windows in HBM.  Products per chunk
// and workgroup size per case (measured: 1024/256 is 11 % faster than
// 2048/512 on R-MAT x B(65536 columns), 2048/512 13 % faster on R-MAT A.A with
// its 2^20-column windows).
constexpr int64_t kLdsBitsCols = int64_t(1) << 18;
constexpr int kCapLarge = 1024;
constexpr int kTpbLarge = 256;
// The count kernel also keeps windows of up to 2^20 columns (128 KB) in LDS,
// one 1024-thread workgroup per CU.
constexpr int64_t kLdsBitsColsWide = int64_t(1) << 20;
constexpr int kTpbLargeWide = 1024;
// The value kernel fits such a bitmap beside 512-product chunk arrays.
constexpr int kCapLargeWideNum = 512;
// Large rows of at least kCountHeavy products among the first
// kCountHeavyRows of the heaviest-first list are counted by one 1024-thread
// workgroup each before the others (narrow bitmaps, number of large rows known).
constexpr int64_t kCountHeavy = int64_t(1) << 18;
constexpr int kCountHeavyRows = 256;
constexpr int kTpbLargeWideNum = 512;
constexpr int kCapLargeHbm = 2048;
constexpr int kTpbLargeHbm = 512;
// Large rows of products whose B has at most kCompactCols columns fold into
// an accumulator in LDS indexed by each column's rank among the row's
// distinct columns, from the presence bitmap the count kernel saves per row
// (kCompactCols / 8 bytes each).
constexpr int64_t kCompactCols = int64_t(1) << 16;
// Partitioned large rows: the count kernel also counts each large row's
// products and distinct columns per window of kPartCols columns; a partition
// kernel moves the row's products (window-local column and product value) to
// HBM scratch, window by window and in serial order inside each window; fold
// tasks of about kPartTaskProducts products then fold each window into an LDS
// accumulator.  Used when the count keeps every column of B in its LDS bitmap
// (at most kPartMaxWin windows) and the scratch fits the device.
#ifndef SPG_PART_LOG
#define SPG_PART_LOG 12
#endif
constexpr int kPartLog = SPG_PART_LOG;
constexpr int64_t kPartCols = int64_t(1) << kPartLog;
constexpr int kPartMaxWin = 1 << (20 - kPartLog);  // C of at most 2^20 columns
constexpr int64_t kPartTaskProducts = int64_t(1) << 15;
constexpr int kTpbPart = kPartMaxWin > 256 ? kPartMaxWin : 256;
constexpr int kCapPart = 1024;
constexpr int kTpbFold = 256;
constexpr int kCapFold = 1024;
constexpr int kCapLargeCompact = 1024;
constexpr int kTpbLargeCompact = 512;
// Single-pass path: rows with at most kFusedCap products and A entries are
// computed by one wave each, kFusedRows consecutive rows per wave, in tiles of
// (kFusedTileTpb / 64) * kFusedRows rows (these are exactly the rows of bins
// W1, W2).  Four-wave tiles compiled for five waves per SIMD (96 VGPRs) keep
// 20 waves per CU: on the 2^20-row 16-per-row product their kernel took 1.515
// ms against 1.543 ms for eight-wave tiles (16 rows, 109 VGPRs, 16 waves per
// CU).  The packed tiles (kFusedTpb) take 16 rows per workgroup.
constexpr int kFusedCap = kCapW2;
#ifndef SPG_FUSED_TPB
#define SPG_FUSED_TPB 512
#endif
constexpr int kFusedTpb = SPG_FUSED_TPB;
#ifndef SPG_FUSED_TILE_TPB
#define SPG_FUSED_TILE_TPB 256
#endif
constexpr int kFusedTileTpb = SPG_FUSED_TILE_TPB;
constexpr int kFusedRows = 2;  // default rows per wave of a tile (1 or 2)
// Dealt tiles (rows_per_wave == kFusedDealRows, a mode id): a tile is 32
// consecutive rows on a 1024-thread workgroup, and its single-pass rows are
// dealt to the 16 waves in row order, two at most per wave.  Where most rows
// are counted ahead (bins W3 and up), a tile's waves then hold about one row
// each, and a wave's two rows are staged together even when not adjacent.
constexpr int kFusedDealRows = 4;
constexpr int kFusedDealTpb = 1024;
constexpr int kFusedDealTile = 32;
// Packed tiles (rows_per_wave == kFusedPackRows, a mode id): the single-pass
// rows of each analyze block (kAnalyzeRows consecutive rows) are cut into
// tiles of kFusedPackCap in row order, at least one tile per block; a tile
// also carries the counts of the block's other rows up to its last
// single-pass row (the block's last tile: to the block's end).  Tiles then
// hold single-pass rows only, however few the block has.
constexpr int kFusedPackRows = 8;
constexpr int kFusedPackCap = 16;
constexpr int kAnalyzeRows = 256;  // rows per analyze block (4 waves of 64)
// Tile t's look-back status is word t * kStatusStride of the status array:
// one status per 64-byte line.  Packed statuses put the few lines that every
// waiting tile polls on one memory channel; spreading them took the 2^20-row
// 16-per-row product's single-pass kernel from 2.06 to 1.77 ms.
#ifndef SPG_STATUS_STRIDE
#define SPG_STATUS_STRIDE 8
#endif
constexpr int kStatusStride = SPG_STATUS_STRIDE;

// Error flags raised by kernels (OR-ed into Control::flags).
enum : unsigned {
    kErrRange = 1u,     // A column outside the B row directory
    kErrMissing = 2u,   // A column names a B row that is neither local nor in the halo
    kErrInternal = 4u,  // capacity/invariant violation (never expected)
    kErrLongRow = 8u    // a referenced B row has more than INT32_MAX entries
};

// Small control block in device memory, mirrored to pinned host memory.
struct Control {
    unsigned long long products;    // sum of P_row
    unsigned long long flags;       // kErr* bits
    unsigned long long actual;      // stored entries after the numeric phase
    unsigned long long structural;  // distinct coordinates (scan total, two-phase path)
    unsigned long long sym_total;   // distinct coordinates counted by the symbolic kernels
    unsigned long long ub_small;    // sum of P_row over rows of bins W1, W2
    unsigned long long reserved;    // entries of the C row layout (single-pass path)
    unsigned long long distinct;    // distinct coordinates of the single-pass rows
    int bin_count[8];               // rows listed per bin (list lengths)
    long long bin_rows[8];          // rows per bin (statistics; listed or not)
    int large_next[3];              // next large row to take: count, HBM-window values, compact values
    int pad_;
    unsigned long long packed_tiles;  // tiles of the packed single-pass kernel
    unsigned long long max_a_row;     // direct single pass: longest row of A
    unsigned long long max_b_row;     // direct single pass: longest row of B
    unsigned long long ntasks;        // partitioned large rows: fold tasks written
    unsigned long long part_products; // partitioned large rows: their products (scratch entries)
    unsigned long long part_max_row;  // partitioned large rows: the largest product count of one
    unsigned long long part_units;    // partitioned large rows: batches of kCapLarge A entries of all large rows
    int part_next;                    // next batch of the partition kernel
    int fold_next;                    // next fold task
    int heavy_units;                  // analyze: pieces of heavy row tiles listed
    int heavy_next;                   // analyze: next piece to take
};

// Per-block partial sums of the analyze pass, reduced by one small kernel
// (a device-scope atomic on one address costs tens of nanoseconds, so the
// hot kernels never add into the control block per wave).
constexpr int kAnalyzeFields = 2 + kNumBins;  // products, ub of bins W1/W2, rows per bin

// A fold task: windows [w0, w1) of large row i (list index).
struct PartTask {
    int64_t s0;    // scratch position of window w0's products
    int64_t out0;  // C position of window w0's first entry (structural layout)
    int32_t i;
    int16_t w0, w1;
};

// State of the partitioned large-row path (large row i of the list at
// i * nwin in the per-window arrays).
struct PartState {
    int64_t *pw;     // [nlarge * nwin] products per window (count kernel)
    int32_t *pd;     // [nlarge * nwin] distinct columns per window (count kernel)
    int64_t *segbase;  // [2 * (nlarge + 1)] batches of kCapLarge A entries per large row, then their exclusive scan
    uint32_t *pseg;    // [batches * nwin] products per window before each batch of its row (count kernel)
    int32_t *pst;    // [nlarge * nwin] stored entries per window (fold kernel)
    int32_t *pgap;   // [nlarge] nonzero when a window of the row dropped exact zeros
    int64_t *poff;   // [2 * (nlarge + 1)] products of each partitioned row, then their exclusive scan
    uint16_t *scol;  // [scratch] window-local columns
    double *sval;    // [scratch] products
    PartTask *task;  // [max_tasks]
    int64_t max_tasks;
    int64_t skip_le;  // rows with snnz <= skip_le belong to the compact kernel (-1: none)
    int nwin;
};

// Outputs of the single-pass kernel.
struct FusedOut {
    int64_t *rowptr;    // [rows+1] C row offsets (all rows)
    int32_t *ccol;
    double *cval;
    uint64_t *status;   // [ntiles] look-back words, zero before the launch
    int64_t *distinct;  // [ntiles] distinct columns of the tile's single-pass rows
    int64_t ntiles;     // tiles launched (packed tiles: an upper bound; Control::packed_tiles is the count)
    int64_t *tmap;      // packed tiles: tile -> (analyze block << 5) | tile of the block
};

// Operands as the kernels see them.  B rows are reached through a directory
// indexed by the global B row id: entries [bstart[k], bend[k]); bstart[k] < 0
// marks a row that is neither local nor in the halo.
struct Operands {
    const int64_t *aptr;
    const int32_t *acol;
    const double *aval;
    int64_t nnz_a;
    int32_t arows;
    int32_t ndir;
    const int64_t *bstart;
    const int64_t *bend;
    const int32_t *bcol;
    const double *bval;
    int32_t bcols;
    int32_t keep_zeros;  // 1: store coordinates whose sum is exactly 0.0 (staged accumulation)
};

// One listed row of a bin: its index, its A-entry count and its first A entry
// (so kernels need not load A's row offsets again).
struct RowRec {
    int32_t row;
    int32_t na;   // A entries (saturated at INT32_MAX; only bins <= kCapBlock rely on it)
    int64_t a0;
};

// Workspace::mode bits.
constexpr int kWsListSmall = 1;  // analyze also lists the rows of bins W1, W2 (two-phase path)
constexpr int kWsNarrow = 2;     // ebst holds 32-bit starts: B has fewer than 2^29 entries (small_b)

struct Workspace {
    int64_t *ebst;        // [nnz(A)] start of the B row each A entry references
    int32_t *elen;        // [nnz(A)] length of that B row (0 for a bad/missing row)
    uint32_t *bdir32;     // [ndir+1] B's row offsets in 32 bits when the directory is B's own
                          // row_ptr and every entry index fits, else null
    int64_t *ub;          // [arows] per-row product count
    RowRec *bin_recs;     // [kNumBins * list_stride] row records per bin
    int64_t list_stride;  // >= arows
    int64_t *snnz;        // [arows] structural count per row
    int64_t *annz;        // [arows] stored count per row
    int64_t *scan_tmp;    // block partials of the scan
    Control *ctrl;
    int64_t *part;        // [analyze blocks * kAnalyzeFields] partial sums
    int mode;             // kWs* bits
    uint32_t *lbits;      // [lblocks * lwidth/32] large-row presence bitmaps (zero between rows)
    double *lacc;         // [lblocks * lwidth] large-row dense accumulators
    int64_t lwidth;       // column window of the large-row path (multiple of 32)
    int lblocks;
};

// Launch helpers.  All are asynchronous on `stream`.
void launch_zero_bytes(void *p, size_t bytes, hipStream_t stream);  // bytes: multiple of 4
void launch_zero_control(Control *ctrl, hipStream_t stream);
// Per-row work bound, bins and bin row lists (analyze.hip).
int64_t analyze_blocks(int32_t rows);
void launch_analyze(const Operands &op, const Workspace &ws, hipStream_t stream);
void launch_dir32(const int64_t *bptr, int64_t n, uint32_t *out, hipStream_t stream);
// The longest rows of A (arows rows) and B (brows rows) into
// ctrl->max_a_row / max_b_row (cleared with the control block), and B's row
// offsets in 32 bits into dir32 (brows + 1 entries; B has fewer than 2^32
// entries).
void launch_rowmax(const int64_t *aptr, int32_t arows, const int64_t *bptr, int32_t brows, uint32_t *dir32,
                   Control *ctrl, hipStream_t stream);
// dst = sum of in[0..n) (one workgroup; scan.hip).
// *dst += sum of in[0, n) (dst is cleared by the caller, e.g. with the control block).
void launch_sum(const int64_t *in, int64_t n, unsigned long long *dst, hipStream_t stream);
// Bin sets of the launchers: bit b selects bin b.
constexpr unsigned kSmallBins = (1u << kBinW1) | (1u << kBinW2);
constexpr unsigned kBigBins = (1u << kBinW3) | (1u << kBinW4) | (1u << kBinBlock) | (1u << kBinLarge);
constexpr unsigned kAllBins = kSmallBins | kBigBins;
// Symbolic phase for the selected bins (symbolic.hip); kernels read their row
// counts from the control block, so no host synchronisation is needed before it.
// With `lsave`, the large-row count also saves each row's presence bitmap
// (row i of the large-row list at word i * lwidth/32; B's columns must fit
// one window) for the compact value kernel.
// With `part` (partition counts: part_available) the large-row count also
// writes each large row's products and distinct columns per window of
// kPartCols columns (pw, pd), the products per window before each batch of
// kCapLarge A entries (pseg), and sums the products and the largest product
// count of the rows with snnz > part->skip_le into the control block.
// nlarge: the number of large rows when the host knows it (-1 otherwise); it
// lets the heaviest be counted by a launch of their own, on `side` beside the
// other rows' count when side, fork and join are given (the side stream is
// forked from and joined back to `stream` by the two events).
void launch_symbolic(const Operands &op, const Workspace &ws, unsigned bins, int cus, hipStream_t stream,
                     uint32_t *lsave = nullptr, const PartState *part = nullptr, int nlarge = -1,
                     hipStream_t side = nullptr, hipEvent_t fork = nullptr, hipEvent_t join = nullptr);
// Batches of kCapLarge A entries of every large row and their exclusive scan
// (ps.segbase; total into ctrl->part_units), before the large-row count
// (large_part.hip).
void launch_part_batches(const Operands &op, const Workspace &ws, const PartState &ps, int nlarge,
                         hipStream_t stream);
// Whether the large rows of a product with `bcols` B columns can be
// partitioned when their count runs over windows of `lwidth` columns: the
// count keeps all of B's columns in its LDS bitmap, in at most kPartMaxWin
// windows.
bool part_available(int64_t bcols, int64_t lwidth);
inline int part_windows(int64_t bcols) { return (int)((bcols + kPartCols - 1) >> kPartLog); }
// Rows with snnz at most this are the compact value kernel's (its LDS
// accumulator holds them in one pass), or -1 when it does not run for windows
// of `lwidth` columns.
int64_t compact_take_le(int64_t lwidth);
// Per-process kernel attributes (dynamic LDS limits); called when a context
// is created, with the device's LDS bytes per workgroup.
hipError_t prepare_kernels(size_t lds_per_block);
bool prepare_numeric_kernels(size_t lds_per_block);  // numeric.hip; false: wide LDS path unavailable
// Single-pass kernel over all rows (fused.hip): computes the rows of bins W1,
// W2 and reserves the symbolic counts of the others.  small_b: every B entry
// index i has 8*i < 2^32.
// rows_per_wave: 1 or 2 (rows of a tile per wave; the tile has kFusedTileTpb/64
// waves), kFusedDealRows (dealt 32-row tiles) or kFusedPackRows (packed
// tiles; fused_tiles does not apply, the plan launches an upper bound).
int64_t fused_tiles(int32_t rows, int rows_per_wave);
// Packed tiles: per analyze block max(1, ceil(single-pass rows / kFusedPackCap))
// from the analyze partials (`part`, `nblocks` blocks) into tcnt, their
// exclusive scan into toff (total into ctrl->packed_tiles; tmp as launch_scan)
// and the tile map (fo.tmap, at most `ntiles` entries).
void launch_pack_tiles(const int64_t *part, int64_t nblocks, int64_t *tcnt, int64_t *toff, int64_t *tmap,
                       int64_t *tmp, Control *ctrl, hipStream_t stream);
void launch_fused(const Operands &op, const Workspace &ws, const FusedOut &fo, bool small_b, int rows_per_wave,
                  hipStream_t stream);
// Direct single pass (no analysis): every row has at most kFusedCap products
// and A entries, which the host checked from the longest rows of A and B.
// Each tile reads its A entries' columns and B's 32-bit row offsets
// (ws.bdir32) itself.  fo.distinct holds kDirectFields x ntiles counters:
// field f of tile t at [f * ntiles + t] (distinct columns, scalar products,
// rows without products, rows of bin W1, rows of bin W2).
constexpr int kDirectFields = 5;
void launch_fused_direct(const Operands &op, const Workspace &ws, const FusedOut &fo, bool small_b,
                         int rows_per_wave, hipStream_t stream);
// Exclusive scan of n int64 counts into out[0..n]; out[n] = total; total also
// stored to *total_dst when non-null (scan.hip).
size_t scan_tmp_elems(int64_t n);
void launch_scan(const int64_t *in, int64_t *out, int64_t n, int64_t *tmp, unsigned long long *total_dst,
                 hipStream_t stream);
// Numeric phase for the selected bins whose host-side counts are nonzero (numeric.hip).
// With `lsave` (bitmaps saved by the count), large rows take the compact
// value kernel, which needs compact_large_available(ws.lwidth).
// The compact kernel runs on `side` when given (forked from and joined back
// to `stream` with the events fork/join), beside the HBM-window kernel: the
// window kernel's heaviest rows bound its duration while its other workgroups
// finish early, and the compact kernel's workgroups fill those CUs.
// With `part`, the large rows that the compact kernel does not take are
// partitioned and folded (launch_part) instead of folded onto HBM windows.
void launch_numeric(const Operands &op, const Workspace &ws, const int *bin_counts, unsigned bins, int cus,
                    const int64_t *crowptr, int32_t *ccol, double *cval, hipStream_t stream,
                    const uint32_t *lsave = nullptr, hipStream_t side = nullptr, hipEvent_t fork = nullptr,
                    hipEvent_t join = nullptr, const PartState *part = nullptr);
// The partitioned large-row path (large_part.hip): scratch offsets of the
// partitioned rows, the partition kernel (which also writes the fold tasks),
// the fold kernel and the move of windows left apart by dropped zeros.
// nlarge: rows in the large-row list.
void launch_part(const Operands &op, const Workspace &ws, const PartState &ps, int nlarge, int cus,
                 const int64_t *crowptr, int32_t *ccol, double *cval, hipStream_t stream);
// Whether the compact large-row value kernel runs for windows of `lwidth`
// columns on this device (its LDS accumulator holds at least one chunk).
bool compact_large_available(int64_t lwidth);
// Copy each row's entries from (rowptr_in) to (rowptr_out); the count of row r
// is rowptr_out[r+1] - rowptr_out[r] (scan.hip).
void launch_compact(int32_t rows, const int64_t *rowptr_in, const int32_t *col_in, const double *val_in,
                    const int64_t *rowptr_out, int32_t *col_out, double *val_out, hipStream_t stream);

}  // namespace spg
