// noalloc.hip -- the drop-in product without any device memory allocation.
//
// For callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops): every intermediate array of a product lives
// in one fixed, file-scope device array, `scratch`, and C goes straight into
// buffers the caller passes in.  Each GPU the library runs on holds its own
// copy of the array from the first context created on it until the process
// ends.  All calls on one GPU share that copy: calls that overlap on one GPU
// race, and everything that must start at zero is cleared by kernels on every
// call.
//
// The pipeline is the single-pass one, reordered so that C comes out compact
// with no compaction buffer:
//   analyze + symbolic counts of the rows counted ahead (bins W3 and up)
//   -> host reads the control block: symbolic bound of C, big rows' entries
//   numeric kernels of the big rows into scratch (their own compact layout)
//   -> their stored counts replace their structural counts as reservations
//   single-pass kernel: small rows computed, offsets of every row, C written
//   -> big rows copied from scratch to their final place
// Results are bit-identical to spg_multiply_begin's.
#include "device_common.hpp"
#include "engine_internal.hpp"

#include <mutex>
#include <string>

// 512 MiB: the workspace of a 2^20-row product with 16 entries per row and
// 256 products per row (about 295 MB, of which 8.4 MB are the
// single-pass tiles' statuses) fits, with room for big rows.
constexpr size_t kScratchBytes = size_t(1) << 29;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace spg {
namespace {

using namespace dev;

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};  // device address of `scratch` per device (after set-up)
bool g_setup[kMaxDevices] = {};
hipError_t g_setup_err[kMaxDevices] = {};

// Window of the large-row accumulators here: 2^16 columns (the presence
// bitmap in LDS), one window per resident large-row workgroup.
constexpr int64_t kNaWindow = int64_t(1) << 16;
constexpr int kNaLargeBlocks = 128;

__device__ __forceinline__ bool counted_ahead(const Operands &op, const Workspace &ws, int64_t r) {
    return classify(ws.ub[r], op.aptr[r + 1] - op.aptr[r]) >= kBinW3;
}

// cnt[r]: structural count of a row counted ahead, else 0.
__global__ void k_na_big_counts(Operands op, Workspace ws, int64_t *cnt) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < op.arows;
         r += (int64_t)gridDim.x * blockDim.x)
        cnt[r] = counted_ahead(op, ws, r) ? ws.snnz[r] : 0;
}

// The big rows' stored counts (written by their numeric kernels into annz)
// become the single-pass kernel's reservations.
__global__ void k_na_reserve(Operands op, Workspace ws) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < op.arows;
         r += (int64_t)gridDim.x * blockDim.x)
        if (counted_ahead(op, ws, r)) ws.snnz[r] = ws.annz[r];
}

// One wave per big row: its stored entries from scratch to C.
__global__ void k_na_move(Operands op, Workspace ws, const int64_t *srow, const int32_t *scol, const double *sval,
                          const int64_t *crow, int32_t *ccol, double *cval) {
    const int lane = lane_id();
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; r < op.arows; r += nw) {
        if (!counted_ahead(op, ws, r)) continue;
        const int64_t n = ws.annz[r], s = srow[r], d = crow[r];
        for (int64_t q = lane; q < n; q += 64) {
            ccol[d + q] = scol[s + q];
            cval[d + q] = sval[s + q];
        }
    }
}

unsigned grid_rows(int64_t rows, int per_block) {
    int64_t g = (rows + per_block - 1) / per_block;
    if (g > 16384) g = 16384;
    return (unsigned)(g < 1 ? 1 : g);
}

// Bump allocation of 256-byte aligned pieces of the scratch array.
struct Carve {
    char *base;
    size_t used = 0;
    size_t need = 0;  // bytes asked for, also past the end
    explicit Carve(void *b) : base(static_cast<char *>(b)) {}
    template <typename T>
    T *take(size_t count) {
        const size_t bytes = (count * sizeof(T) + 255) & ~size_t(255);
        need += bytes;
        if (used + bytes > kScratchBytes) return nullptr;
        T *p = reinterpret_cast<T *>(base + used);
        used += bytes;
        return p;
    }
    bool ok() const { return need <= kScratchBytes; }
};

// The control block (in scratch) into the context's pinned mirror.
spg_status read_scratch_control(spg_context *ctx, const Control *ctrl) {
    SPG_HIP(hipMemcpyAsync(ctx->h_ctrl, ctrl, sizeof(Control), hipMemcpyDeviceToHost, ctx->stream),
            "read control block");
    SPG_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return SPG_OK;
}

spg_status too_big(spg_context *ctx, const Carve &c, const char *what) {
    return set_err(ctx, SPG_ENOMEM,
                   std::string("allocation-free product: ") + what + " need " + std::to_string(c.need) +
                       " bytes of scratch; the maximum is " + std::to_string(kScratchBytes) + " bytes");
}

}  // namespace

// Called by spg_create with `device` current: the device address of this
// device's copy of `scratch` (host code never uses the array's own name).
void noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_setup[device]) return;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        ptr = nullptr;
    }
    g_scratch[device] = ptr;
    g_setup_err[device] = e;
    g_setup[device] = true;
}

}  // namespace spg

using namespace spg;

extern "C" spg_status spg_multiply_noalloc(spg_context *ctx, const spg_csr_view *A, const spg_csr_view *B,
                                           int64_t *row_ptr, int32_t *col_idx, double *values, int64_t capacity,
                                           int64_t *nnz_c) {
    if (!ctx) return SPG_EINVAL;
    if (!A || !B || !nnz_c || !row_ptr) return set_err(ctx, SPG_EINVAL, "null argument");
    *nnz_c = 0;
    ctx->have_result = false;
    if (A->location != SPG_DEVICE || B->location != SPG_DEVICE)
        return set_err(ctx, SPG_EINVAL, "the allocation-free product takes device views of A and B");
    if (capacity < 0 || (capacity > 0 && (!col_idx || !values)))
        return set_err(ctx, SPG_EINVAL, "null output buffer");
    spg_status s;
    if ((s = check_host_csr(ctx, A, "A")) != SPG_OK) return s;
    if ((s = check_host_csr(ctx, B, "B_local")) != SPG_OK) return s;
    SPG_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    void *base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        if (ctx->device >= 0 && ctx->device < kMaxDevices && g_setup[ctx->device]) base = g_scratch[ctx->device];
        if (!base) {
            const hipError_t e = ctx->device >= 0 && ctx->device < kMaxDevices ? g_setup_err[ctx->device]
                                                                                : hipErrorInvalidDevice;
            return set_err(ctx, SPG_EDEVICE,
                           std::string("scratch array unavailable on this device: ") + hipGetErrorString(e));
        }
    }
    hipStream_t st = ctx->stream;
    const int32_t rows = A->rows;
    const size_t R = rows > 0 ? (size_t)rows : 1;
    const size_t E = A->nnz > 0 ? (size_t)A->nnz : 1;

    ProductPlan p;
    Operands &op = p.op;
    op.aptr = A->row_ptr;
    op.acol = A->col_idx;
    op.aval = A->values;
    op.nnz_a = A->nnz;
    op.arows = rows;
    op.ndir = B->rows;
    op.bstart = B->row_ptr;
    op.bend = B->row_ptr + 1;
    op.bcol = B->col_idx;
    op.bval = B->values;
    op.bcols = B->cols;
    p.rows = rows;
    p.cols = B->cols;
    p.small_b = narrow_b(ctx, B->nnz);
    p.b_span = B->nnz;
    p.fused = true;
    p.fused_rows = ctx->fused_rows;

    // Workspace of the counting phase.
    Carve cv(base);
    Workspace &ws = p.ws;
    ws = Workspace{};
    ws.mode = p.small_b ? kWsNarrow : 0;  // single pass: only the big rows are listed
    ws.ctrl = cv.take<Control>(1);
    ws.part = cv.take<int64_t>((size_t)(analyze_blocks(rows) + 1) * kAnalyzeFields);
    ws.ub = cv.take<int64_t>(R);
    ws.snnz = cv.take<int64_t>(R);
    ws.annz = cv.take<int64_t>(R);
    ws.bin_recs = cv.take<RowRec>(R * kNumBins);
    ws.list_stride = (int64_t)R;
    ws.scan_tmp = cv.take<int64_t>(scan_tmp_elems(rows));
    ws.ebst = cv.take<int64_t>(p.small_b ? (E + 1) / 2 : E);  // 4-byte starts when narrow
    ws.elen = cv.take<int32_t>(E);
    int64_t width = ((int64_t)p.cols + 31) / 32 * 32;
    if (width > kNaWindow) width = kNaWindow;
    if (width < 32) width = 32;
    ws.lwidth = width;
    ws.lblocks = ctx->cus * 2 < kNaLargeBlocks ? ctx->cus * 2 : kNaLargeBlocks;
    ws.lbits = cv.take<uint32_t>((size_t)ws.lblocks * (size_t)(width / 32));
    if (!cv.ok()) return too_big(ctx, cv, "the counting phase's arrays");

    launch_zero_bytes(ws.lbits, (size_t)ws.lblocks * (size_t)(width / 32) * 4, st);
    launch_zero_control(ws.ctrl, st);
    launch_analyze(op, ws, st);
    launch_symbolic(op, ws, kBigBins, ctx->cus, st);
    if ((s = read_scratch_control(ctx, ws.ctrl)) != SPG_OK) return s;
    Control h = *ctx->h_ctrl;
    if ((s = flag_error(ctx, h.flags)) != SPG_OK) {
        if (s == SPG_ERANGE) return set_err(ctx, SPG_EMISSING_ROW, "missing a remote sparse row required by A");
        return s;
    }
    const int64_t bound = (int64_t)(h.ub_small + h.sym_total);
    if (bound > capacity) {
        *nnz_c = bound;
        return set_err(ctx, SPG_ENOMEM,
                       "output capacity " + std::to_string(capacity) + " is below the symbolic bound " +
                           std::to_string(bound) + " of C");
    }
    for (int b = 0; b < 8; ++b) p.bin_count[b] = h.bin_count[b];

    // Workspace of the values phase: the big rows' own layout, the large-row
    // windows, the single-pass kernel's look-back state.
    const int64_t nbig = (int64_t)h.sym_total;
    int64_t *srow = cv.take<int64_t>(R + 1);
    int32_t *scol = cv.take<int32_t>(nbig > 0 ? (size_t)nbig : 1);
    double *sval = cv.take<double>(nbig > 0 ? (size_t)nbig : 1);
    if (h.bin_count[kBinLarge] > 0) ws.lacc = cv.take<double>((size_t)ws.lblocks * (size_t)width);
    p.fo = FusedOut{};
    if (p.fused_rows == kFusedPackRows) {
        p.pack_blocks = analyze_blocks(rows);
        const int64_t nf = (int64_t)h.bin_rows[kBinW1] + (int64_t)h.bin_rows[kBinW2];
        p.fo.ntiles = rows > 0 ? p.pack_blocks + (nf + kFusedPackCap - 1) / kFusedPackCap : 0;
        const size_t nb = (size_t)(p.pack_blocks > 0 ? p.pack_blocks : 1);
        p.pack_cnt = cv.take<int64_t>(nb);
        p.pack_off = cv.take<int64_t>(nb);
        p.fo.tmap = cv.take<int64_t>((size_t)(p.fo.ntiles > 0 ? p.fo.ntiles : 1));
    } else {
        p.fo.ntiles = fused_tiles(rows, p.fused_rows);
    }
    const size_t nt = (size_t)(p.fo.ntiles > 0 ? p.fo.ntiles : 1);
    p.fo.status = cv.take<uint64_t>(nt * kStatusStride);
    p.fo.distinct = cv.take<int64_t>(nt);
    if (!cv.ok()) return too_big(ctx, cv, "the values phase's arrays (big rows, look-back state)");
    p.fo.rowptr = row_ptr;
    p.fo.ccol = col_idx;
    p.fo.cval = values;

    // Big rows first, into their own compact layout in scratch.
    const bool big = h.bin_count[kBinW3] + h.bin_count[kBinW4] + h.bin_count[kBinBlock] + h.bin_count[kBinLarge] > 0;
    if (big) {
        hipLaunchKernelGGL(k_na_big_counts, dim3(grid_rows(rows, 256)), dim3(256), 0, st, op, ws, ws.annz);
        launch_scan(ws.annz, srow, rows, ws.scan_tmp, nullptr, st);
        launch_numeric(op, ws, p.bin_count, kBigBins, ctx->cus, srow, scol, sval, st);
        hipLaunchKernelGGL(k_na_reserve, dim3(grid_rows(rows, 256)), dim3(256), 0, st, op, ws);
    }
    // Single pass over every row: C's offsets (big rows at their stored
    // counts) and the small rows' entries, straight into the caller's buffers.
    if (p.fused_rows == kFusedPackRows)
        launch_pack_tiles(ws.part, p.pack_blocks, p.pack_cnt, p.pack_off, p.fo.tmap, ws.scan_tmp, ws.ctrl, st);
    launch_zero_bytes(p.fo.status, nt * kStatusStride * 8, st);
    if (rows == 0) launch_zero_bytes(row_ptr, 8, st);
    launch_fused(op, ws, p.fo, p.small_b, p.fused_rows, st);
    if (p.fo.ntiles > 0) launch_sum(p.fo.distinct, p.fo.ntiles, &ws.ctrl->distinct, st);
    if (big)
        hipLaunchKernelGGL(k_na_move, dim3(grid_rows((int64_t)rows * 64, 256)), dim3(256), 0, st, op, ws, srow, scol,
                           sval, row_ptr, col_idx, values);
    SPG_HIP(hipGetLastError(), "kernel launch");
    if ((s = read_scratch_control(ctx, ws.ctrl)) != SPG_OK) return s;
    h = *ctx->h_ctrl;
    if ((s = flag_error(ctx, h.flags)) != SPG_OK) return s;
    const int64_t nnz = rows > 0 ? (int64_t)h.reserved : 0;
    if (nnz > bound) return set_err(ctx, SPG_EDEVICE, "internal capacity invariant violated");
    spg_stats stats{};
    stats.products = (int64_t)h.products;
    for (int b = 0; b < kNumBins; ++b) stats.rows_per_bin[b] = h.bin_rows[b];
    stats.bins = kNumBins;
    stats.structural_nnz = (int64_t)(h.distinct + h.sym_total);
    stats.nnz = nnz;
    stats.single_pass = 1;
    stats.wide_format = p.small_b ? 0 : 1;
    ctx->stats = stats;
    *nnz_c = nnz;
    return SPG_OK;
}
