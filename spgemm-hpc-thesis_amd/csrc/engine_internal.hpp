// engine_internal.hpp -- the context object and the product pipeline shared by
// the drop-in entry points (engine.hip) and the batched ones (batch.hip).
#pragma once

#include "internal.hpp"
#include "spgemm_amd.h"

#include <string>

namespace spg {

// Growth-only device buffer (no allocation once a context is warm).  With
// `headroom` (batch buffers, rebuilt for products of varying sizes), a buffer
// that has to grow again takes 1/8 more than asked, so that slowly growing
// sizes reallocate rarely: freeing and allocating tens of GB stalls the
// device for about a second.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    bool zero_on_grow = false;
    bool headroom = false;
    // Bytes held after ensure(want).
    size_t size_for(size_t want) const {
        if (want <= bytes && p) return bytes;
        size_t n = p && headroom ? want + want / 8 : want;
        n = n < 256 ? 256 : n;
        return (n + 255) & ~size_t(255);
    }
    hipError_t ensure(size_t want, hipStream_t stream) {
        if (want <= bytes && p) return hipSuccess;
        const size_t n = size_for(want);
        if (p) {
            hipError_t e = hipStreamSynchronize(stream);
            if (e != hipSuccess) return e;
            hipFree(p);
            p = nullptr;
            bytes = 0;
        }
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            return e;
        }
        bytes = n;
        if (zero_on_grow) launch_zero_bytes(p, n, stream);
        return hipSuccess;
    }
    void release() {
        if (p) hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <typename T>
    T *as() const { return static_cast<T *>(p); }
};

// Buffers a phase needs, with their sizes: what ensure() on each would add to
// the device bytes held (a growing buffer is freed before it is allocated
// again).
struct BufferNeeds {
    static constexpr int kMax = 32;
    int n = 0;
    DevBuf *buf[kMax] = {};
    size_t bytes[kMax] = {};
    const char *what[kMax] = {};
    void add(DevBuf &b, size_t sz, const char *w) {
        buf[n] = &b;
        bytes[n] = sz;
        what[n] = w;
        ++n;
    }
    uint64_t growth() const {
        uint64_t g = 0;
        for (int i = 0; i < n; ++i) {
            const uint64_t after = buf[i]->size_for(bytes[i]);
            if (after > buf[i]->bytes) g += after - buf[i]->bytes;
        }
        return g;
    }
};

// Everything one product needs on the device, fixed once its counts are
// known.  A plan can be replayed (eagerly or from a captured HIP graph) for
// operands of the same structure without host synchronisation.
struct ProductPlan {
    Operands op{};
    Workspace ws{};
    int32_t rows = 0, cols = 0;
    bool small_b = false;   // every B entry's byte offset fits 32 bits (entries < 2^29)
    int64_t b_span = 0;     // B entries reachable through the directory
    bool fused = false;  // single-pass kernel for bins W1, W2
    int fused_rows = kFusedRows;
    int bin_count[8] = {};
    int64_t capacity = 0;  // C entries allocated
    FusedOut fo{};
    int64_t pack_blocks = 0;             // packed tiles: analyze blocks
    int64_t *pack_cnt = nullptr;         // packed tiles: tiles per block, then their offsets
    int64_t *pack_off = nullptr;
    bool compact = false;                // large rows take the compact value kernel (B columns <= kCompactCols)
    uint32_t *lsave = nullptr;           // compact: the large rows' saved presence bitmaps, sized for this plan
    bool part_counts = false;            // the large-row count writes per-window counts (part_available)
    bool part = false;                   // the large rows the compact kernel leaves are partitioned
    PartState ps{};                      // their state (pw, pd from the count; the rest for the values)
};

}  // namespace spg

struct spg_context {
    int device = 0;
    int cus = 256;
    size_t total_mem = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    bool timing = false;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    hipStream_t side = nullptr;                      // second stream of the large-row kernels
    hipEvent_t fork = nullptr, join = nullptr;       // its fork and join (timing disabled)
    std::string err;
    spg::Control *h_ctrl = nullptr;  // pinned mirror
    // uploaded operands (host inputs) and the halo directory
    spg::DevBuf a_ptr, a_col, a_val, b_ptr, b_col, b_val, dir_start, dir_end;
    // workspace
    spg::DevBuf ebst, elen, ub, bin_rows, snnz, annz, scan_tmp, ctrl, lbits, lacc, tile_status, tile_distinct, part, dir32;
    spg::DevBuf halo_err, halo_cnt, halo_tmp;  // device halo directory and packing
    spg::DevBuf pack_cnt, pack_off, pack_map;  // packed single-pass tiles
    spg::DevBuf lsave;                          // large rows' presence bitmaps (compact value kernel)
    spg::DevBuf ppw, ppd, ppst, ppgap, ppoff, pscol, psval, ptask, psegb, pseg;  // partitioned large rows
    int path = SPG_PATH_AUTO;
    int fused_rows = spg::kFusedRows;
    bool force_wide = false;  // SPG_FORMAT_WIDE: 64-bit entry stream and B addressing whatever B's size
    bool allow_direct = true; // products whose rows all fit the single pass skip the analysis (not for batch plans)
    int large_mode = SPG_LARGE_AUTO;  // SPG_LARGE_WINDOWS: no partitioned large rows
    int64_t lbits_width = 0;
    // results (ping-pong for compaction)
    spg::DevBuf rowptr[2], ccol[2], cval[2];
    int cur = 0;
    bool have_result = false;
    int32_t res_rows = 0, res_cols = 0;
    int64_t res_nnz = 0;
    spg_stats stats{};
    // Every device buffer the context owns.
    template <typename F>
    void for_each_buffer(F f) {
        spg::DevBuf *bufs[] = {&lsave, &a_ptr, &a_col, &a_val, &b_ptr, &b_col, &b_val, &dir_start, &dir_end,
                               &ebst, &elen, &ub, &bin_rows, &snnz, &annz, &scan_tmp, &ctrl, &lbits, &lacc,
                               &tile_status, &tile_distinct, &part, &dir32, &halo_err, &halo_cnt, &halo_tmp,
                               &pack_cnt, &pack_off, &pack_map, &rowptr[0], &rowptr[1], &ccol[0], &ccol[1],
                               &cval[0], &cval[1], &ppw, &ppd, &ppst, &ppgap, &ppoff, &pscol, &psval, &ptask,
                               &psegb, &pseg};
        for (spg::DevBuf *b : bufs) f(*b);
    }
    uint64_t device_bytes() {
        uint64_t n = 0;
        for_each_buffer([&](spg::DevBuf &b) { n += b.bytes; });
        return n;
    }
};

namespace spg {

// Whether a product whose B rows hold b_span entries takes the 32-bit entry
// stream and B addressing (every B entry index i then has 8*i < 2^32), unless
// the context pins the 64-bit format.
inline bool narrow_b(const spg_context *ctx, int64_t b_span) {
    return !ctx->force_wide && b_span < (int64_t(1) << 29);
}

spg_status set_err(spg_context *c, spg_status s, const std::string &msg);
spg_status hip_err(spg_context *c, hipError_t e, const char *what);
spg_status check_host_csr(spg_context *ctx, const spg_csr_view *m, const char *name);
// Operands of a product whose local B block and halo rows are device views:
// entries [B_local | halo] and the global-row directory built on the device
// (halo.hip), covering global rows [0, ndir).  b_span receives the number of B
// entries.
spg_status device_halo_operands(spg_context *ctx, const spg_csr_view *A, const spg_csr_view *B, int32_t b_first_row,
                                const spg_halo_view *halo, int32_t ndir, Operands &op, int64_t &b_span);
spg_status read_control(spg_context *ctx);
// Device address of the allocation-free entry point's scratch array on
// `device` (the current device), looked up once per device (noalloc.hip).
void noalloc_setup(int device);
spg_status flag_error(spg_context *ctx, unsigned long long flags);

#define SPG_HIP(call, what)                                       \
    do {                                                          \
        hipError_t e_ = (call);                                   \
        if (e_ != hipSuccess) return spg::hip_err(ctx, e_, what); \
    } while (0)

// The count workspace of a product with `rows` A rows, nnz_a A entries, a
// directory of ndir B rows (the 32-bit copy when dir32) and `cols` C columns:
// its buffers with their sizes, and the large-row windows' width and count.
void workspace_needs(spg_context *ctx, int32_t rows, int64_t nnz_a, int32_t ndir, bool dir32, int64_t cols,
                     BufferNeeds &v, int64_t &lwidth, int &lblocks);
// Workspace buffers for p.op / p.rows / p.cols; fills p.ws.
spg_status plan_workspace(spg_context *ctx, ProductPlan &p);
// Zero the control block, analyze, symbolic kernels of `bins` (and the scan
// of all structural counts into C's row offsets when `scan`).
void enqueue_counts(spg_context *ctx, const ProductPlan &p, unsigned bins, bool scan);
// C buffers and the single-pass outputs for a plan whose counts are in h.
spg_status plan_values(spg_context *ctx, ProductPlan &p, const Control &h);
// Single-pass kernel and numeric kernels of the plan.
void enqueue_values(spg_context *ctx, const ProductPlan &p);
// Compaction when exact zeros left gaps; publishes the result on the context.
spg_status finish_product(spg_context *ctx, const ProductPlan &p, const Control &h, spg_stats &stats);
// The count phase of a product: workspace, analyze and the symbolic kernels,
// up to the point where C's size is known (the counts are left in h).  `done`
// when the direct single pass has already produced the result.
spg_status count_product(spg_context *ctx, ProductPlan &p, Control &h, bool &done);
// Device bytes the value phase of a counted plan adds to the context's
// buffers: C's entries at the plan's capacity (exact structural counts on the
// two-phase path; the small rows' work bounds plus the big rows' exact counts
// on the single pass), the large rows' accumulator windows and the
// single-pass tile state, less what the buffers already hold.  Compaction
// buffers (results with exact zeros) are not included.
uint64_t value_bytes(spg_context *ctx, const ProductPlan &p, const Control &h);
// The value phase of a counted plan: C buffers, value kernels, compaction.
spg_status value_product(spg_context *ctx, ProductPlan &p, const Control &counted);
// The whole product with its two host synchronisations (count_product, then
// value_product); leaves the plan filled for replays.
spg_status run_product(spg_context *ctx, ProductPlan &p);

}  // namespace spg
