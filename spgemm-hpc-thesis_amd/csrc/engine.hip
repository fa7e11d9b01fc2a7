// engine.hip -- host side of the MI355X SpGEMM engine and the C ABI of
// include/spgemm_amd.h.
//
// The context owns growth-only device buffers (no allocation in steady state),
// a stream, and a pinned mirror of the kernels' control block.  A multiply is:
//   analyze (per-row work bound, bins, bin row lists), symbolic phase (exact
//   distinct-column counts), scan -> C row_ptr      -> sync 1: errors, nnz(C)
//   numeric phase                                   -> sync 2: stored entries
//   compaction, only when exact zeros were dropped.
#include "engine_internal.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <unordered_set>
#include <vector>

using namespace spg;

namespace spg {

spg_status set_err(spg_context *c, spg_status s, const std::string &msg) {
    if (c) c->err = msg;
    return s;
}

spg_status hip_err(spg_context *c, hipError_t e, const char *what) {
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    (void)hipGetLastError();
    if (e == hipErrorOutOfMemory) return set_err(c, SPG_ENOMEM, m);
    return set_err(c, SPG_EDEVICE, m);
}

spg_status check_host_csr(spg_context *ctx, const spg_csr_view *m, const char *name) {
    if (m->rows < 0 || m->cols < 0 || m->nnz < 0) return set_err(ctx, SPG_EINVAL, std::string(name) + ": negative dimension");
    if (!m->row_ptr) return set_err(ctx, SPG_EINVAL, std::string(name) + ": null row_ptr");
    if (m->nnz > 0 && (!m->col_idx || !m->values))
        return set_err(ctx, SPG_EINVAL, std::string(name) + ": null column/value arrays");
    if (m->location != SPG_HOST) return SPG_OK;  // device views are trusted (documented)
    if (m->row_ptr[0] != 0 || m->row_ptr[m->rows] != m->nnz)
        return set_err(ctx, SPG_EINVAL, std::string(name) + ": row_ptr must start at 0 and end at nnz");
    for (int32_t r = 0; r < m->rows; ++r)
        if (m->row_ptr[r + 1] < m->row_ptr[r])
            return set_err(ctx, SPG_EINVAL, std::string(name) + ": row_ptr is not nondecreasing");
    return SPG_OK;
}

spg_status read_control(spg_context *ctx) {
    SPG_HIP(hipMemcpyAsync(ctx->h_ctrl, ctx->ctrl.p, sizeof(Control), hipMemcpyDeviceToHost, ctx->stream),
            "read control block");
    SPG_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return SPG_OK;
}

spg_status flag_error(spg_context *ctx, unsigned long long flags) {
    if (flags & kErrLongRow) return set_err(ctx, SPG_EOVERFLOW, "a B row has more than INT32_MAX entries");
    if (flags & kErrMissing)
        return set_err(ctx, SPG_EMISSING_ROW, "missing a remote sparse row required by A");
    if (flags & kErrRange)
        return set_err(ctx, SPG_ERANGE, "A contains a column index outside B's row range");
    if (flags & kErrInternal) return set_err(ctx, SPG_EDEVICE, "internal capacity invariant violated");
    return SPG_OK;
}

void workspace_needs(spg_context *ctx, int32_t rows, int64_t nnz_a, int32_t ndir, bool dir32, int64_t cols,
                     BufferNeeds &v, int64_t &lwidth, int &lblocks) {
    const size_t R = rows > 0 ? (size_t)rows : 1;
    const size_t E = nnz_a > 0 ? (size_t)nnz_a : 1;
    v.add(ctx->ub, R * 8, "hipMalloc(ub)");
    v.add(ctx->ebst, E * 8, "hipMalloc(entry starts)");
    v.add(ctx->elen, E * 4, "hipMalloc(entry lengths)");
    v.add(ctx->bin_rows, R * sizeof(RowRec) * kNumBins, "hipMalloc(bin lists)");
    v.add(ctx->snnz, R * 8, "hipMalloc(snnz)");
    v.add(ctx->annz, R * 8, "hipMalloc(annz)");
    v.add(ctx->scan_tmp, scan_tmp_elems(rows) * 8, "hipMalloc(scan)");
    v.add(ctx->ctrl, sizeof(Control), "hipMalloc(ctrl)");
    v.add(ctx->part, (size_t)(analyze_blocks(rows) + 1) * kAnalyzeFields * 8, "hipMalloc(partials)");
    v.add(ctx->rowptr[0], ((size_t)rows + 1) * 8, "hipMalloc(row_ptr C)");
    if (dir32) v.add(ctx->dir32, ((size_t)ndir + 1) * 4, "hipMalloc(B directory)");
    // Large-row path: one column window and one presence bitmap per resident
    // workgroup (kept all-zero between rows); the accumulators are allocated
    // only when large rows exist.
    int64_t width = (cols + 31) / 32 * 32;
    const int64_t kMaxWindow = int64_t(1) << 22;
    if (width > kMaxWindow) width = kMaxWindow;
    if (width < 32) width = 32;
    const size_t per_block = (size_t)width * 8 + (size_t)width / 8;
    size_t budget = ctx->total_mem / 16;
    if (budget > (size_t(4) << 30)) budget = size_t(4) << 30;
    size_t nb = budget / per_block;
    const size_t cap = (size_t)ctx->cus * 2;
    if (nb > cap) nb = cap;
    if (nb < 1) nb = 1;
    v.add(ctx->lbits, nb * (size_t)(width / 32) * 4, "hipMalloc(large bitmaps)");
    lwidth = width;
    lblocks = (int)nb;
}

spg_status plan_workspace(spg_context *ctx, ProductPlan &p) {
    hipStream_t st = ctx->stream;
    // B's row_ptr as the directory (bend = bstart + 1) with 32-bit entry
    // indices: analyze gathers a 4-byte copy of it (half the cache lines).
    const bool dir32 = p.b_span < (int64_t(1) << 32) && p.op.ndir > 0 && p.op.bend == p.op.bstart + 1;
    BufferNeeds v;
    int64_t width = 0;
    int nb = 0;
    workspace_needs(ctx, p.rows, p.op.nnz_a, p.op.ndir, dir32, p.cols, v, width, nb);
    ctx->lbits.zero_on_grow = true;
    if (ctx->lbits_width != width) {  // bitmaps are laid out per window width
        ctx->lbits.release();
        ctx->lbits_width = width;
    }
    for (int i = 0; i < v.n; ++i) SPG_HIP(v.buf[i]->ensure(v.bytes[i], st), v.what[i]);

    Workspace &ws = p.ws;
    ws = Workspace{};
    ws.ebst = ctx->ebst.as<int64_t>();
    // Every start is at most b_span (missing rows write 0).  When B is small
    // (b_span < 2^29, the plan's small_b, which also picks the single-pass
    // kernel's 32-bit addressing) the stream holds 32-bit starts: 4 bytes per
    // A entry less written by analyze and read by every later kernel.
    if (p.small_b) ws.mode |= kWsNarrow;
    ws.elen = ctx->elen.as<int32_t>();
    ws.ub = ctx->ub.as<int64_t>();
    ws.bin_recs = ctx->bin_rows.as<RowRec>();
    ws.list_stride = (int64_t)(p.rows > 0 ? p.rows : 1);
    ws.snnz = ctx->snnz.as<int64_t>();
    ws.annz = ctx->annz.as<int64_t>();
    ws.scan_tmp = ctx->scan_tmp.as<int64_t>();
    ws.ctrl = ctx->ctrl.as<Control>();
    ws.part = ctx->part.as<int64_t>();
    if (dir32) ws.bdir32 = ctx->dir32.as<uint32_t>();
    ws.lwidth = width;
    ws.lblocks = nb;
    ws.lbits = ctx->lbits.as<uint32_t>();
    ws.lacc = ctx->lacc.as<double>();
    return SPG_OK;
}

void enqueue_counts(spg_context *ctx, const ProductPlan &p, unsigned bins, bool scan) {
    hipStream_t st = ctx->stream;
    launch_zero_control(p.ws.ctrl, st);
    if (p.ws.bdir32) launch_dir32(p.op.bstart, (int64_t)p.op.ndir + 1, p.ws.bdir32, st);
    launch_analyze(p.op, p.ws, st);
    // replays of a plan whose large rows are partitioned recompute the
    // batches of the large rows (the control block was cleared)
    const bool part = p.part_counts && (bins & (1u << kBinLarge)) && p.bin_count[kBinLarge] > 0 && p.ps.pseg;
    if (part) launch_part_batches(p.op, p.ws, p.ps, p.bin_count[kBinLarge], st);
    launch_symbolic(p.op, p.ws, bins, ctx->cus, st, p.lsave, part ? &p.ps : nullptr);
    if (scan) launch_scan(p.ws.snnz, ctx->rowptr[0].as<int64_t>(), p.rows, p.ws.scan_tmp, &p.ws.ctrl->structural, st);
}

// The counting phase of a plan whose large-row count writes per-row arrays
// sized by the number of large rows (the compact value kernel's saved
// bitmaps, the partition's per-window counts), which is known only after the
// other bins are counted: the large rows are counted after a first read of the
// control block (then C's row offsets are scanned when `scan`).  Leaves the
// control block's current state in h.
spg_status counts_deferred(spg_context *ctx, ProductPlan &p, unsigned bins, bool scan, Control &h) {
    hipStream_t st = ctx->stream;
    p.lsave = nullptr;
    const bool part_counts = p.part_counts;
    p.part_counts = false;  // the first pass counts no large rows
    enqueue_counts(ctx, p, bins & ~(1u << kBinLarge), false);
    p.part_counts = part_counts;
    spg_status s;
    if ((s = read_control(ctx)) != SPG_OK) return s;
    h = *ctx->h_ctrl;
    if ((s = flag_error(ctx, h.flags)) != SPG_OK) return s;
    const int nlarge = (bins & (1u << kBinLarge)) ? h.bin_count[kBinLarge] : 0;
    if (nlarge == 0 && !scan) return SPG_OK;
    if (nlarge > 0) {
        if (p.compact) {
            SPG_HIP(ctx->lsave.ensure((size_t)nlarge * (size_t)(p.ws.lwidth / 8), st), "hipMalloc(large-row bitmaps)");
            p.lsave = ctx->lsave.as<uint32_t>();
        }
        if (p.part_counts) {
            // per large row and window, and per batch of kCapLarge A entries
            // (at most one per large row plus one per kCapLarge A entries)
            const size_t cells = (size_t)nlarge * (size_t)p.ps.nwin;
            const size_t batches = (size_t)nlarge + (size_t)(p.op.nnz_a / kCapLarge) + 1;
            SPG_HIP(ctx->ppw.ensure(cells * 8, st), "hipMalloc(window products)");
            SPG_HIP(ctx->ppd.ensure(cells * 4, st), "hipMalloc(window columns)");
            SPG_HIP(ctx->psegb.ensure(2 * ((size_t)nlarge + 1) * 8, st), "hipMalloc(batch offsets)");
            SPG_HIP(ctx->pseg.ensure(batches * (size_t)p.ps.nwin * 4, st), "hipMalloc(batch window counts)");
            p.ps.pw = ctx->ppw.as<int64_t>();
            p.ps.pd = ctx->ppd.as<int32_t>();
            p.ps.segbase = ctx->psegb.as<int64_t>();
            p.ps.pseg = ctx->pseg.as<uint32_t>();
            launch_part_batches(p.op, p.ws, p.ps, nlarge, st);
        }
        launch_symbolic(p.op, p.ws, 1u << kBinLarge, ctx->cus, st, p.lsave, p.part_counts ? &p.ps : nullptr, nlarge,
                        ctx->side, ctx->fork, ctx->join);
    }
    if (scan) launch_scan(p.ws.snnz, ctx->rowptr[0].as<int64_t>(), p.rows, p.ws.scan_tmp, &p.ws.ctrl->structural, st);
    if ((s = read_control(ctx)) != SPG_OK) return s;
    h = *ctx->h_ctrl;
    return SPG_OK;
}

namespace {

// The buffers the value phase of a counted plan needs, with their sizes.
struct ValueBuffers : BufferNeeds {
    int64_t capacity = 0, ntiles = 0, pack_blocks = 0, max_tasks = 0;
};

ValueBuffers value_buffers(spg_context *ctx, const ProductPlan &p, const Control &h) {
    ValueBuffers v;
    v.capacity = p.fused ? (int64_t)(h.ub_small + h.sym_total) : (int64_t)h.structural;
    v.add(ctx->ccol[0], (size_t)v.capacity * 4, "hipMalloc(C col)");
    v.add(ctx->cval[0], (size_t)v.capacity * 8, "hipMalloc(C val)");
    const int nlarge = h.bin_count[kBinLarge];
    if (nlarge > 0 && p.part) {
        // partitioned large rows: per-window stored counts, gap flags,
        // scratch offsets, the scratch and the fold tasks (every task but a
        // row's last holds at least kPartTaskProducts products)
        const size_t cells = (size_t)nlarge * (size_t)p.ps.nwin;
        const size_t P = (size_t)h.part_products;
        v.max_tasks = (int64_t)nlarge + (int64_t)(P / (size_t)kPartTaskProducts) + 1;
        v.add(ctx->ppst, cells * 4, "hipMalloc(window stored counts)");
        v.add(ctx->ppgap, (size_t)nlarge * 4, "hipMalloc(window gaps)");
        v.add(ctx->ppoff, 2 * ((size_t)nlarge + 1) * 8, "hipMalloc(scratch offsets)");
        v.add(ctx->pscol, P * 2, "hipMalloc(scratch columns)");
        v.add(ctx->psval, P * 8, "hipMalloc(scratch products)");
        v.add(ctx->ptask, (size_t)v.max_tasks * sizeof(PartTask), "hipMalloc(fold tasks)");
    } else if (nlarge > 0) {
        v.add(ctx->lacc, (size_t)p.ws.lblocks * (size_t)p.ws.lwidth * 8, "hipMalloc(large accumulators)");
    }
    if (p.fused) {
        if (p.fused_rows == kFusedPackRows) {
            // Packed tiles: at most one per block plus one per kFusedPackCap
            // single-pass rows; the device counts them exactly.
            v.pack_blocks = analyze_blocks(p.rows);
            const int64_t nf = (int64_t)h.bin_rows[kBinW1] + (int64_t)h.bin_rows[kBinW2];
            v.ntiles = p.rows > 0 ? v.pack_blocks + (nf + kFusedPackCap - 1) / kFusedPackCap : 0;
            const size_t nb = (size_t)(v.pack_blocks > 0 ? v.pack_blocks : 1);
            v.add(ctx->pack_cnt, nb * 8, "hipMalloc(packed tile counts)");
            v.add(ctx->pack_off, nb * 8, "hipMalloc(packed tile offsets)");
            v.add(ctx->pack_map, (size_t)(v.ntiles > 0 ? v.ntiles : 1) * 8, "hipMalloc(packed tile map)");
        } else {
            v.ntiles = fused_tiles(p.rows, p.fused_rows);
        }
        const size_t nt = (size_t)(v.ntiles > 0 ? v.ntiles : 1);
        v.add(ctx->tile_status, nt * kStatusStride * 8, "hipMalloc(tile status)");
        v.add(ctx->tile_distinct, nt * 8, "hipMalloc(tile counts)");
    }
    return v;
}

}  // namespace

uint64_t value_bytes(spg_context *ctx, const ProductPlan &p, const Control &h) {
    return value_buffers(ctx, p, h).growth();
}

spg_status plan_values(spg_context *ctx, ProductPlan &p, const Control &h) {
    hipStream_t st = ctx->stream;
    for (int b = 0; b < 8; ++b) p.bin_count[b] = h.bin_count[b];
    const ValueBuffers v = value_buffers(ctx, p, h);
    for (int i = 0; i < v.n; ++i) SPG_HIP(v.buf[i]->ensure(v.bytes[i], st), v.what[i]);
    p.capacity = v.capacity;
    if (h.bin_count[kBinLarge] > 0 && !p.part) p.ws.lacc = ctx->lacc.as<double>();
    if (h.bin_count[kBinLarge] > 0 && p.part) {
        p.ps.pst = ctx->ppst.as<int32_t>();
        p.ps.pgap = ctx->ppgap.as<int32_t>();
        p.ps.poff = ctx->ppoff.as<int64_t>();
        p.ps.scol = ctx->pscol.as<uint16_t>();
        p.ps.sval = ctx->psval.as<double>();
        p.ps.task = ctx->ptask.as<PartTask>();
        p.ps.max_tasks = v.max_tasks;
    }
    p.fo = FusedOut{};
    if (p.fused) {
        p.fo.ntiles = v.ntiles;
        if (p.fused_rows == kFusedPackRows) {
            p.pack_blocks = v.pack_blocks;
            p.pack_cnt = ctx->pack_cnt.as<int64_t>();
            p.pack_off = ctx->pack_off.as<int64_t>();
            p.fo.tmap = ctx->pack_map.as<int64_t>();
        }
        p.fo.rowptr = ctx->rowptr[0].as<int64_t>();
        p.fo.ccol = ctx->ccol[0].as<int32_t>();
        p.fo.cval = ctx->cval[0].as<double>();
        p.fo.status = ctx->tile_status.as<uint64_t>();
        p.fo.distinct = ctx->tile_distinct.as<int64_t>();
    }
    return SPG_OK;
}

void enqueue_values(spg_context *ctx, const ProductPlan &p) {
    hipStream_t st = ctx->stream;
    if (p.fused) {
        if (p.fused_rows == kFusedPackRows)
            launch_pack_tiles(p.ws.part, p.pack_blocks, p.pack_cnt, p.pack_off, p.fo.tmap, p.ws.scan_tmp, p.ws.ctrl,
                              st);
        launch_zero_bytes(p.fo.status, (size_t)(p.fo.ntiles > 0 ? p.fo.ntiles : 1) * kStatusStride * 8, st);
        if (p.rows == 0) launch_zero_bytes(p.fo.rowptr, 8, st);
        launch_fused(p.op, p.ws, p.fo, p.small_b, p.fused_rows, st);
    }
    launch_numeric(p.op, p.ws, p.bin_count, p.fused ? kBigBins : kAllBins, ctx->cus, ctx->rowptr[0].as<int64_t>(),
                   ctx->ccol[0].as<int32_t>(), ctx->cval[0].as<double>(), st, p.lsave, ctx->side, ctx->fork, ctx->join,
                   p.part ? &p.ps : nullptr);
    if (p.fused && p.fo.ntiles > 0) launch_sum(p.fo.distinct, p.fo.ntiles, &p.ws.ctrl->distinct, st);
}

spg_status finish_product(spg_context *ctx, const ProductPlan &p, const Control &h, spg_stats &stats) {
    hipStream_t st = ctx->stream;
    spg_status s;
    if ((s = flag_error(ctx, h.flags)) != SPG_OK) return s;
    // Entries of the row layout: structural counts (two-phase path), or the
    // single-pass rows' stored counts plus the big rows' structural counts.
    // The numeric kernels report the entries they store (ctrl.actual); the
    // single-pass rows' stored total is the layout minus the big rows' share.
    const int64_t layout = p.fused ? (int64_t)h.reserved : (int64_t)h.structural;
    stats.products = (int64_t)h.products;
    for (int b = 0; b < kNumBins; ++b) stats.rows_per_bin[b] = h.bin_rows[b];
    stats.bins = kNumBins;
    stats.structural_nnz = p.fused ? (int64_t)(h.distinct + h.sym_total) : (int64_t)h.structural;
    stats.single_pass = p.fused ? 1 : 0;
    stats.wide_format = p.small_b ? 0 : 1;
    stats.partitioned = p.part && h.bin_count[kBinLarge] > 0 ? 1 : 0;
    const int64_t nnz = p.fused ? (int64_t)(h.reserved - h.sym_total + h.actual) : (int64_t)h.actual;
    int result = 0;
    if (nnz != layout) {
        // Exact zeros were dropped: pack rows to their final offsets.
        const int o2 = 1;
        SPG_HIP(ctx->rowptr[o2].ensure(((size_t)p.rows + 1) * 8, st), "hipMalloc(row_ptr C2)");
        SPG_HIP(ctx->ccol[o2].ensure((size_t)nnz * 4, st), "hipMalloc(C2 col)");
        SPG_HIP(ctx->cval[o2].ensure((size_t)nnz * 8, st), "hipMalloc(C2 val)");
        launch_scan(p.ws.annz, ctx->rowptr[o2].as<int64_t>(), p.rows, p.ws.scan_tmp, nullptr, st);
        launch_compact(p.rows, ctx->rowptr[0].as<int64_t>(), ctx->ccol[0].as<int32_t>(), ctx->cval[0].as<double>(),
                       ctx->rowptr[o2].as<int64_t>(), ctx->ccol[o2].as<int32_t>(), ctx->cval[o2].as<double>(), st);
        SPG_HIP(hipStreamSynchronize(st), "hipStreamSynchronize");
        result = o2;
        stats.compacted = 1;
    }
    SPG_HIP(hipGetLastError(), "kernel launch");
    stats.nnz = nnz;
    ctx->cur = result;
    ctx->res_rows = p.rows;
    ctx->res_cols = p.cols;
    ctx->res_nnz = nnz;
    ctx->have_result = true;
    ctx->stats = stats;
    return SPG_OK;
}

// Direct single pass.  When A's longest row has at most kFusedCap entries
// and its product with B's longest row is at most kFusedCap, every row of C
// has at most kFusedCap products: the single-pass kernel then reads each
// entry's column and B's 32-bit row offsets itself, and the per-row analysis,
// its entry stream and the count kernels do not run.  One small kernel finds
// the longest rows (and writes the 32-bit offsets); otherwise *done stays
// false and the regular pipeline runs.  Results are those of the regular
// pipeline.  Taken by SPG_PATH_AUTO for A of at most kDirectMaxEntries
// entries: there the launches saved outweigh the single-pass kernel's extra
// dependent load (column, then B's row offsets) per entry.  Measured with
// uniform operands, 16 or 8 entries per row: 16384 rows 0.098 against 0.109
// ms, 65536 rows -3 %; from 2^18 rows x 16 on the two are within 2 %.
constexpr int64_t kDirectMaxEntries = int64_t(1) << 20;

spg_status try_direct(spg_context *ctx, ProductPlan &p, bool &done) {
    done = false;
    if (!ctx->allow_direct || p.rows <= 0 || !p.ws.bdir32 || p.op.nnz_a > kDirectMaxEntries) return SPG_OK;
    if (ctx->path != SPG_PATH_AUTO || p.fused_rows != kFusedRows) return SPG_OK;
    hipStream_t st = ctx->stream;
    spg_status s;
    if (ctx->timing) SPG_HIP(hipEventRecord(ctx->ev[0], st), "hipEventRecord");
    launch_zero_control(p.ws.ctrl, st);
    launch_rowmax(p.op.aptr, p.rows, p.op.bstart, p.op.ndir, p.ws.bdir32, p.ws.ctrl, st);
    if (ctx->timing) SPG_HIP(hipEventRecord(ctx->ev[1], st), "hipEventRecord");
    if ((s = read_control(ctx)) != SPG_OK) return s;
    Control h = *ctx->h_ctrl;
    const unsigned long long ma = h.max_a_row, mb = h.max_b_row;
    if (ma > (unsigned long long)kFusedCap || ma * mb > (unsigned long long)kFusedCap) return SPG_OK;
    // C capacity: every row's products bound its entries.
    const unsigned long long by_entries = (unsigned long long)p.op.nnz_a * mb;
    const unsigned long long by_rows = (unsigned long long)p.rows * ma * mb;
    p.capacity = (int64_t)(by_entries < by_rows ? by_entries : by_rows);
    if (p.capacity < 1) p.capacity = 1;
    SPG_HIP(ctx->ccol[0].ensure((size_t)p.capacity * 4, st), "hipMalloc(C col)");
    SPG_HIP(ctx->cval[0].ensure((size_t)p.capacity * 8, st), "hipMalloc(C val)");
    p.fused = true;
    p.fo = FusedOut{};
    p.fo.ntiles = fused_tiles(p.rows, p.fused_rows);
    const size_t nt = (size_t)p.fo.ntiles;
    SPG_HIP(ctx->tile_status.ensure(nt * kStatusStride * 8, st), "hipMalloc(tile status)");
    SPG_HIP(ctx->tile_distinct.ensure(nt * kDirectFields * 8, st), "hipMalloc(tile counts)");
    p.fo.rowptr = ctx->rowptr[0].as<int64_t>();
    p.fo.ccol = ctx->ccol[0].as<int32_t>();
    p.fo.cval = ctx->cval[0].as<double>();
    p.fo.status = ctx->tile_status.as<uint64_t>();
    p.fo.distinct = ctx->tile_distinct.as<int64_t>();
    if (ctx->timing) SPG_HIP(hipEventRecord(ctx->ev[2], st), "hipEventRecord");
    launch_zero_bytes(p.fo.status, nt * kStatusStride * 8, st);
    launch_fused_direct(p.op, p.ws, p.fo, p.small_b, p.fused_rows, st);
    Control *dc = p.ws.ctrl;
    unsigned long long *dst[kDirectFields] = {&dc->distinct, &dc->products,
                                              reinterpret_cast<unsigned long long *>(&dc->bin_rows[kBinEmpty]),
                                              reinterpret_cast<unsigned long long *>(&dc->bin_rows[kBinW1]),
                                              reinterpret_cast<unsigned long long *>(&dc->bin_rows[kBinW2])};
    for (int f = 0; f < kDirectFields; ++f) launch_sum(p.fo.distinct + (size_t)f * nt, p.fo.ntiles, dst[f], st);
    if (ctx->timing) SPG_HIP(hipEventRecord(ctx->ev[3], st), "hipEventRecord");
    SPG_HIP(hipGetLastError(), "kernel launch");
    if ((s = read_control(ctx)) != SPG_OK) return s;
    h = *ctx->h_ctrl;
    if ((s = flag_error(ctx, h.flags)) != SPG_OK) return s;
    const int64_t nnz = (int64_t)h.reserved;
    if (nnz > p.capacity) return set_err(ctx, SPG_EDEVICE, "internal capacity invariant violated");
    spg_stats stats{};
    stats.products = (int64_t)h.products;
    for (int b = 0; b < kNumBins; ++b) stats.rows_per_bin[b] = h.bin_rows[b];
    stats.bins = kNumBins;
    stats.structural_nnz = (int64_t)h.distinct;
    stats.nnz = nnz;
    stats.single_pass = 1;
    stats.wide_format = p.small_b ? 0 : 1;
    stats.direct = 1;
    if (ctx->timing) {
        float t = 0;
        SPG_HIP(hipEventElapsedTime(&t, ctx->ev[0], ctx->ev[3]), "hipEventElapsedTime");
        stats.ms_total = t;
        SPG_HIP(hipEventElapsedTime(&t, ctx->ev[0], ctx->ev[1]), "hipEventElapsedTime");
        stats.ms_symbolic = t;
        SPG_HIP(hipEventElapsedTime(&t, ctx->ev[2], ctx->ev[3]), "hipEventElapsedTime");
        stats.ms_numeric = t;
    }
    ctx->cur = 0;
    ctx->res_rows = p.rows;
    ctx->res_cols = p.cols;
    ctx->res_nnz = nnz;
    ctx->have_result = true;
    ctx->stats = stats;
    done = true;
    return SPG_OK;
}

spg_status count_product(spg_context *ctx, ProductPlan &p, Control &h, bool &done) {
    hipStream_t st = ctx->stream;
    ctx->have_result = false;
    spg_status s;
    p.fused_rows = ctx->fused_rows;
    if ((s = plan_workspace(ctx, p)) != SPG_OK) return s;
    done = false;
    if ((s = try_direct(ctx, p, done)) != SPG_OK || done) return s;
    if (ctx->timing) SPG_HIP(hipEventRecord(ctx->ev[0], st), "hipEventRecord");
    // Single-pass path: only the rows of the big bins are listed and counted
    // ahead; the rows of bins W1, W2 are computed in one pass by the fused kernel.
    const bool try_fused = ctx->path != SPG_PATH_TWO_PHASE;
    p.ws.mode = (p.ws.mode & ~kWsListSmall) | (try_fused ? 0 : kWsListSmall);
    // Large rows of products whose B columns fit the compact kernel's window
    // fold in LDS; their count then saves their bitmaps.
    p.compact = compact_large_available(p.ws.lwidth) && (int64_t)p.cols <= p.ws.lwidth;
    p.lsave = nullptr;
    // Large rows the compact kernel leaves are partitioned by column window
    // when the count can count each window; the value phase decides once the
    // scratch size is known.
    p.part = false;
    p.ps = PartState{};
    p.ps.skip_le = p.compact ? compact_take_le(p.ws.lwidth) : -1;
    p.part_counts = ctx->large_mode == SPG_LARGE_AUTO && part_available(p.op.bcols, p.ws.lwidth);
    p.ps.nwin = p.part_counts ? part_windows(p.op.bcols) : 0;
    h = Control{};
    if (p.compact || p.part_counts) {
        if ((s = counts_deferred(ctx, p, try_fused ? kBigBins : kAllBins, !try_fused, h)) != SPG_OK) return s;
        if (ctx->timing) SPG_HIP(hipEventRecord(ctx->ev[1], st), "hipEventRecord");
    } else {
        enqueue_counts(ctx, p, try_fused ? kBigBins : kAllBins, !try_fused);
        if (ctx->timing) SPG_HIP(hipEventRecord(ctx->ev[1], st), "hipEventRecord");
        if ((s = read_control(ctx)) != SPG_OK) return s;
        h = *ctx->h_ctrl;
    }
    if ((s = flag_error(ctx, h.flags)) != SPG_OK) return s;
    // C capacity of the single-pass path: the small rows' work bounds plus the
    // big rows' exact counts.  Taken when it fits half of the device memory
    // (always when the caller asked for it).
    p.fused = false;
    if (try_fused) {
        const uint64_t fused_cap = h.ub_small + h.sym_total;
        p.fused = ctx->path == SPG_PATH_SINGLE_PASS || fused_cap * 12ull <= (uint64_t)(ctx->total_mem / 2);
        if (!p.fused) {
            // Two-phase after all: list the small rows too (analyze again) and
            // count them.
            p.ws.mode |= kWsListSmall;
            if (p.compact || p.part_counts) {
                if ((s = counts_deferred(ctx, p, kAllBins, true, h)) != SPG_OK) return s;
            } else {
                enqueue_counts(ctx, p, kAllBins, true);
                if ((s = read_control(ctx)) != SPG_OK) return s;
                h = *ctx->h_ctrl;
            }
            if ((s = flag_error(ctx, h.flags)) != SPG_OK) return s;
        }
    }
    // The large rows the compact kernel leaves are partitioned when their
    // scratch (10 B per product) fits a quarter of the device and no row has
    // 2^31 products or more (the count's window counters are 32-bit).
    if (p.part_counts && h.bin_count[kBinLarge] > 0 && h.part_products > 0)
        p.part = h.part_max_row < (1ull << 31) && 10ull * h.part_products <= (uint64_t)(ctx->total_mem / 4);
    return SPG_OK;
}

spg_status value_product(spg_context *ctx, ProductPlan &p, const Control &counted) {
    hipStream_t st = ctx->stream;
    spg_stats stats{};
    spg_status s;
    if ((s = plan_values(ctx, p, counted)) != SPG_OK) return s;
    if (ctx->timing) SPG_HIP(hipEventRecord(ctx->ev[2], st), "hipEventRecord");
    enqueue_values(ctx, p);
    if (ctx->timing) SPG_HIP(hipEventRecord(ctx->ev[3], st), "hipEventRecord");
    if ((s = read_control(ctx)) != SPG_OK) return s;
    const Control h = *ctx->h_ctrl;
    if ((s = finish_product(ctx, p, h, stats)) != SPG_OK) return s;
    if (ctx->timing) {
        float t = 0;
        SPG_HIP(hipEventElapsedTime(&t, ctx->ev[0], ctx->ev[3]), "hipEventElapsedTime");
        ctx->stats.ms_total = t;
        SPG_HIP(hipEventElapsedTime(&t, ctx->ev[0], ctx->ev[1]), "hipEventElapsedTime");
        ctx->stats.ms_symbolic = t;
        SPG_HIP(hipEventElapsedTime(&t, ctx->ev[2], ctx->ev[3]), "hipEventElapsedTime");
        ctx->stats.ms_numeric = t;
    }
    return SPG_OK;
}

// The whole product on the device.  On success the result is
// rowptr[cur], ccol[cur], cval[cur] with res_nnz entries.
spg_status run_product(spg_context *ctx, ProductPlan &p) {
    Control h{};
    bool done = false;
    spg_status s = count_product(ctx, p, h, done);
    if (s != SPG_OK || done) return s;
    return value_product(ctx, p, h);
}

}  // namespace spg

namespace {

// Device pointers of a view, uploading host data into the given buffers.
spg_status stage(spg_context *ctx, const spg_csr_view *m, DevBuf &p, DevBuf &c, DevBuf &v, const int64_t *&dp,
                 const int32_t *&dc, const double *&dv) {
    if (m->location == SPG_DEVICE) {
        dp = m->row_ptr;
        dc = m->col_idx;
        dv = m->values;
        return SPG_OK;
    }
    const size_t rp = ((size_t)m->rows + 1) * sizeof(int64_t);
    SPG_HIP(p.ensure(rp, ctx->stream), "hipMalloc(row_ptr)");
    SPG_HIP(c.ensure((size_t)m->nnz * sizeof(int32_t), ctx->stream), "hipMalloc(col)");
    SPG_HIP(v.ensure((size_t)m->nnz * sizeof(double), ctx->stream), "hipMalloc(val)");
    SPG_HIP(hipMemcpyAsync(p.p, m->row_ptr, rp, hipMemcpyHostToDevice, ctx->stream), "upload row_ptr");
    if (m->nnz) {
        SPG_HIP(hipMemcpyAsync(c.p, m->col_idx, (size_t)m->nnz * 4, hipMemcpyHostToDevice, ctx->stream), "upload col");
        SPG_HIP(hipMemcpyAsync(v.p, m->values, (size_t)m->nnz * 8, hipMemcpyHostToDevice, ctx->stream), "upload val");
    }
    dp = p.as<int64_t>();
    dc = c.as<int32_t>();
    dv = v.as<double>();
    return SPG_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char *spg_status_string(spg_status s) {
    switch (s) {
        case SPG_OK: return "ok";
        case SPG_EINVAL: return "invalid argument";
        case SPG_EOVERFLOW: return "count overflow";
        case SPG_ENOMEM: return "out of memory";
        case SPG_EMISSING_ROW: return "missing remote B row";
        case SPG_ERANGE: return "column index outside B's row range";
        case SPG_EDEVICE: return "device error";
        case SPG_ENODEVICE: return "no HIP device";
        case SPG_ESTATE: return "invalid call sequence";
        case SPG_EPARSE: return "parse error";
        case SPG_EBUSY: return "busy";
    }
    return "unknown status";
}

const char *spg_last_error(const spg_context *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

spg_status spg_create(spg_context **out, int device) {
    if (!out) return SPG_EINVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        return SPG_ENODEVICE;
    }
    if (device < 0 || device >= n) return SPG_EINVAL;
    spg_context *ctx = new (std::nothrow) spg_context();
    if (!ctx) return SPG_ENOMEM;
    ctx->device = device;
    hipError_t e = hipSetDevice(device);
    hipDeviceProp_t prop{};
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, device);
    if (e == hipSuccess) e = prepare_kernels(prop.sharedMemPerBlock);
    if (e == hipSuccess) noalloc_setup(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipHostMalloc((void **)&ctx->h_ctrl, sizeof(Control), hipHostMallocDefault);
    for (int i = 0; e == hipSuccess && i < 4; ++i) e = hipEventCreate(&ctx->ev[i]);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->join, hipEventDisableTiming);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        spg_destroy(ctx);
        return SPG_EDEVICE;
    }
    ctx->cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    ctx->total_mem = prop.totalGlobalMem;
    ctx->stream = ctx->own_stream;
    *out = ctx;
    return SPG_OK;
}

spg_status spg_destroy(spg_context *ctx) {
    if (!ctx) return SPG_OK;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ctx->for_each_buffer([](DevBuf &b) { b.release(); });
    for (hipEvent_t &e : ctx->ev)
        if (e) hipEventDestroy(e);
    if (ctx->h_ctrl) hipHostFree(ctx->h_ctrl);
    if (ctx->fork) hipEventDestroy(ctx->fork);
    if (ctx->join) hipEventDestroy(ctx->join);
    if (ctx->side) hipStreamDestroy(ctx->side);
    if (ctx->own_stream) hipStreamDestroy(ctx->own_stream);
    delete ctx;
    return SPG_OK;
}

spg_status spg_set_stream(spg_context *ctx, void *stream) {
    if (!ctx) return SPG_EINVAL;
    ctx->stream = stream ? static_cast<hipStream_t>(stream) : ctx->own_stream;
    return SPG_OK;
}

spg_status spg_context_stream(const spg_context *ctx, void **stream) {
    if (!ctx || !stream) return SPG_EINVAL;
    *stream = ctx->stream;
    return SPG_OK;
}

spg_status spg_context_bind(spg_context *ctx) {
    if (!ctx) return SPG_EINVAL;
    SPG_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    return SPG_OK;
}

spg_status spg_set_path(spg_context *ctx, int path) {
    if (!ctx) return SPG_EINVAL;
    if (path < SPG_PATH_AUTO || path > SPG_PATH_SINGLE_PASS_PACKED) return set_err(ctx, SPG_EINVAL, "unknown path");
    const bool dealt = path == SPG_PATH_SINGLE_PASS_DEALT, packed = path == SPG_PATH_SINGLE_PASS_PACKED;
    ctx->path = dealt || packed ? SPG_PATH_SINGLE_PASS : path;
    ctx->fused_rows = dealt ? kFusedDealRows : packed ? kFusedPackRows : kFusedRows;
    return SPG_OK;
}

spg_status spg_set_format(spg_context *ctx, int format) {
    if (!ctx) return SPG_EINVAL;
    if (format != SPG_FORMAT_AUTO && format != SPG_FORMAT_WIDE) return set_err(ctx, SPG_EINVAL, "unknown format");
    ctx->force_wide = format == SPG_FORMAT_WIDE;
    return SPG_OK;
}

spg_status spg_set_large(spg_context *ctx, int mode) {
    if (!ctx) return SPG_EINVAL;
    if (mode != SPG_LARGE_AUTO && mode != SPG_LARGE_WINDOWS) return set_err(ctx, SPG_EINVAL, "unknown large-row mode");
    ctx->large_mode = mode;
    return SPG_OK;
}

spg_status spg_set_timing(spg_context *ctx, int enabled) {
    if (!ctx) return SPG_EINVAL;
    ctx->timing = enabled != 0;
    return SPG_OK;
}

spg_status spg_last_stats(const spg_context *ctx, spg_stats *out) {
    if (!ctx || !out) return SPG_EINVAL;
    *out = ctx->stats;
    return SPG_OK;
}

namespace {

// C rows [row_begin, row_end) of A times B (the whole product for [0, A->rows)).
spg_status multiply_rows(spg_context *ctx, const spg_csr_view *A, const spg_csr_view *B, int32_t b_first_row,
                         const spg_halo_view *halo, int32_t row_begin, int32_t row_end, int64_t *nnz_c) {
    if (!ctx) return SPG_EINVAL;
    if (!A || !B || !nnz_c) return set_err(ctx, SPG_EINVAL, "null argument");
    ctx->have_result = false;
    SPG_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    spg_status s;
    if ((s = check_host_csr(ctx, A, "A")) != SPG_OK) return s;
    if ((s = check_host_csr(ctx, B, "B_local")) != SPG_OK) return s;
    // The reference's spgemm() compares no shapes: A's columns are resolved
    // row by row, and one that is neither local nor in the halo is fail()ed
    // (reported below as SPG_EMISSING_ROW).
    const bool local_only = halo == nullptr && b_first_row == 0;
    if (b_first_row < 0 || (int64_t)b_first_row + B->rows > INT32_MAX)
        return set_err(ctx, SPG_EINVAL, "local B block outside the int32 row range");
    if (row_begin < 0 || row_end < row_begin || row_end > A->rows)
        return set_err(ctx, SPG_EINVAL, "row range outside A");
    // Global B rows the directory covers: A's column range and the local block.
    const int32_t ndir = std::max(A->cols, b_first_row + B->rows);

    Operands op{};
    op.arows = row_end - row_begin;
    op.nnz_a = A->nnz;
    op.bcols = B->cols;
    if ((s = stage(ctx, A, ctx->a_ptr, ctx->a_col, ctx->a_val, op.aptr, op.acol, op.aval)) != SPG_OK) return s;
    op.aptr += row_begin;  // entries keep their absolute positions
    const int64_t *bptr = nullptr;
    const int32_t *bcol = nullptr;
    const double *bval = nullptr;
    int64_t b_span = B->nnz;
    if (local_only) {
        if ((s = stage(ctx, B, ctx->b_ptr, ctx->b_col, ctx->b_val, bptr, bcol, bval)) != SPG_OK) return s;
        op.ndir = B->rows;
        op.bstart = bptr;
        op.bend = bptr + 1;
        op.bcol = bcol;
        op.bval = bval;
    } else if (B->location == SPG_DEVICE || (halo && halo->location == SPG_DEVICE)) {
        // Device-resident local block and halo: directory built on the device.
        if (B->location != SPG_DEVICE || (halo && halo->location != SPG_DEVICE))
            return set_err(ctx, SPG_EINVAL, "B_local and the halo must both be host or both device views");
        if ((s = device_halo_operands(ctx, A, B, b_first_row, halo, ndir, op, b_span)) != SPG_OK) return s;
    } else {
        // Directory over global B rows: local block first, then halo slots
        // (bRowForColumn order, spgemm_exchange.cpp:75-87).  Entries are
        // concatenated on the device: [B_local entries | halo entries].
        if (halo && (halo->n_rows < 0 || (halo->n_rows > 0 && (!halo->global_rows || !halo->row_ptr))))
            return set_err(ctx, SPG_EINVAL, "malformed halo view");
        const int64_t hn = halo ? halo->n_rows : 0;
        const int64_t hnnz = hn ? halo->row_ptr[hn] : 0;
        b_span = B->nnz + hnnz;
        std::vector<int64_t> ds((size_t)ndir, -1), de((size_t)ndir, -1);
        for (int32_t r = 0; r < B->rows; ++r) {
            ds[(size_t)b_first_row + r] = B->row_ptr[r];
            de[(size_t)b_first_row + r] = B->row_ptr[r + 1];
        }
        std::unordered_set<int32_t> seen;
        for (int64_t slot = 0; slot < hn; ++slot) {
            const int32_t g = halo->global_rows[slot];
            if (g < 0 || g >= ndir) return set_err(ctx, SPG_EINVAL, "halo row outside A's column range");
            if (!seen.insert(g).second) return set_err(ctx, SPG_EINVAL, "duplicate halo row");
            if (halo->row_ptr[slot + 1] < halo->row_ptr[slot] || halo->row_ptr[slot] < 0)
                return set_err(ctx, SPG_EINVAL, "halo row_ptr is negative or not nondecreasing");
            if (g >= b_first_row && g < b_first_row + B->rows) continue;  // local rows take precedence
            ds[(size_t)g] = B->nnz + halo->row_ptr[slot];
            de[(size_t)g] = B->nnz + halo->row_ptr[slot + 1];
        }
        const size_t tot = (size_t)(B->nnz + hnnz);
        SPG_HIP(ctx->b_col.ensure(tot * 4, ctx->stream), "hipMalloc(B col)");
        SPG_HIP(ctx->b_val.ensure(tot * 8, ctx->stream), "hipMalloc(B val)");
        SPG_HIP(ctx->dir_start.ensure((size_t)ndir * 8, ctx->stream), "hipMalloc(dir)");
        SPG_HIP(ctx->dir_end.ensure((size_t)ndir * 8, ctx->stream), "hipMalloc(dir)");
        if (B->nnz) {
            SPG_HIP(hipMemcpyAsync(ctx->b_col.p, B->col_idx, (size_t)B->nnz * 4, hipMemcpyHostToDevice, ctx->stream), "upload B col");
            SPG_HIP(hipMemcpyAsync(ctx->b_val.p, B->values, (size_t)B->nnz * 8, hipMemcpyHostToDevice, ctx->stream), "upload B val");
        }
        if (hnnz) {
            SPG_HIP(hipMemcpyAsync(ctx->b_col.as<int32_t>() + B->nnz, halo->col_idx, (size_t)hnnz * 4,
                                   hipMemcpyHostToDevice, ctx->stream), "upload halo col");
            SPG_HIP(hipMemcpyAsync(ctx->b_val.as<double>() + B->nnz, halo->values, (size_t)hnnz * 8,
                                   hipMemcpyHostToDevice, ctx->stream), "upload halo val");
        }
        if (ndir) {
            SPG_HIP(hipMemcpyAsync(ctx->dir_start.p, ds.data(), ds.size() * 8, hipMemcpyHostToDevice, ctx->stream), "upload dir");
            SPG_HIP(hipMemcpyAsync(ctx->dir_end.p, de.data(), de.size() * 8, hipMemcpyHostToDevice, ctx->stream), "upload dir");
        }
        SPG_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");  // host vectors go out of scope
        op.ndir = ndir;
        op.bstart = ctx->dir_start.as<int64_t>();
        op.bend = ctx->dir_end.as<int64_t>();
        op.bcol = ctx->b_col.as<int32_t>();
        op.bval = ctx->b_val.as<double>();
    }
    // Every B entry index i reachable through the directory has 8*i < 2^32:
    // the single-pass kernel then addresses B with 32-bit offsets.
    ProductPlan plan;
    plan.op = op;
    plan.rows = op.arows;
    plan.cols = B->cols;
    plan.small_b = narrow_b(ctx, b_span);
    plan.b_span = b_span;
    s = run_product(ctx, plan);
    // bRowForColumn (spgemm_exchange.cpp:75-87): a column outside the
    // directory names a row that is neither local nor in the halo.
    if (s == SPG_ERANGE) return set_err(ctx, SPG_EMISSING_ROW, "missing a remote sparse row required by A");
    if (s != SPG_OK) return s;
    *nnz_c = ctx->res_nnz;
    return SPG_OK;
}

}  // namespace

spg_status spg_multiply_begin(spg_context *ctx, const spg_csr_view *A, const spg_csr_view *B, int32_t b_first_row,
                              const spg_halo_view *halo, int64_t *nnz_c) {
    if (!ctx) return SPG_EINVAL;
    if (!A) return set_err(ctx, SPG_EINVAL, "null argument");
    return multiply_rows(ctx, A, B, b_first_row, halo, 0, A->rows, nnz_c);
}

spg_status spg_multiply_rows(spg_context *ctx, const spg_csr_view *A, const spg_csr_view *B, int32_t b_first_row,
                             const spg_halo_view *halo, int32_t row_begin, int32_t row_end, int64_t *nnz_c) {
    return multiply_rows(ctx, A, B, b_first_row, halo, row_begin, row_end, nnz_c);
}

spg_status spg_multiply_finish(spg_context *ctx, int64_t *row_ptr, int32_t *col_idx, double *values) {
    if (!ctx) return SPG_EINVAL;
    if (!ctx->have_result) return set_err(ctx, SPG_ESTATE, "no product to finish");
    if (!row_ptr || (ctx->res_nnz > 0 && (!col_idx || !values))) return set_err(ctx, SPG_EINVAL, "null output buffer");
    const int c = ctx->cur;
    SPG_HIP(hipMemcpyAsync(row_ptr, ctx->rowptr[c].p, ((size_t)ctx->res_rows + 1) * 8, hipMemcpyDeviceToHost, ctx->stream),
            "download row_ptr");
    if (ctx->res_nnz) {
        SPG_HIP(hipMemcpyAsync(col_idx, ctx->ccol[c].p, (size_t)ctx->res_nnz * 4, hipMemcpyDeviceToHost, ctx->stream),
                "download col");
        SPG_HIP(hipMemcpyAsync(values, ctx->cval[c].p, (size_t)ctx->res_nnz * 8, hipMemcpyDeviceToHost, ctx->stream),
                "download val");
    }
    SPG_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return SPG_OK;
}

spg_status spg_result_rows(spg_context *ctx, int32_t row_begin, int32_t row_end, int64_t *row_ptr, int32_t *col_idx,
                           double *values, int64_t capacity, int64_t *nnz_rows) {
    if (!ctx) return SPG_EINVAL;
    if (!ctx->have_result) return set_err(ctx, SPG_ESTATE, "no product available");
    if (!row_ptr || !nnz_rows || row_begin < 0 || row_end < row_begin || row_end > ctx->res_rows)
        return set_err(ctx, SPG_EINVAL, "bad row range or null buffer");
    const int c = ctx->cur;
    const size_t n = (size_t)(row_end - row_begin) + 1;
    SPG_HIP(hipMemcpyAsync(row_ptr, ctx->rowptr[c].as<int64_t>() + row_begin, n * 8, hipMemcpyDeviceToHost, ctx->stream),
            "download row_ptr");
    SPG_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    const int64_t base = row_ptr[0];
    const int64_t cnt = row_ptr[n - 1] - base;
    *nnz_rows = cnt;
    if (cnt > capacity) return set_err(ctx, SPG_EINVAL, "output buffers too small for the row range");
    if (cnt > 0) {
        if (!col_idx || !values) return set_err(ctx, SPG_EINVAL, "null output buffer");
        SPG_HIP(hipMemcpyAsync(col_idx, ctx->ccol[c].as<int32_t>() + base, (size_t)cnt * 4, hipMemcpyDeviceToHost,
                               ctx->stream), "download col");
        SPG_HIP(hipMemcpyAsync(values, ctx->cval[c].as<double>() + base, (size_t)cnt * 8, hipMemcpyDeviceToHost,
                               ctx->stream), "download val");
        SPG_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    }
    for (size_t i = 0; i < n; ++i) row_ptr[i] -= base;
    return SPG_OK;
}

spg_status spg_result_device(spg_context *ctx, spg_csr_view *out) {
    if (!ctx || !out) return SPG_EINVAL;
    if (!ctx->have_result) return set_err(ctx, SPG_ESTATE, "no product available");
    const int c = ctx->cur;
    out->rows = ctx->res_rows;
    out->cols = ctx->res_cols;
    out->nnz = ctx->res_nnz;
    out->row_ptr = ctx->rowptr[c].as<int64_t>();
    out->col_idx = ctx->ccol[c].as<int32_t>();
    out->values = ctx->cval[c].as<double>();
    out->location = SPG_DEVICE;
    return SPG_OK;
}

struct spg_device_matrix {
    int device = 0;
    int32_t rows = 0, cols = 0;
    int64_t nnz = 0;
    void *row_ptr = nullptr, *col = nullptr, *val = nullptr;
};

spg_status spg_matrix_upload(spg_context *ctx, const spg_csr_view *host, spg_device_matrix **out) {
    if (!ctx) return SPG_EINVAL;
    if (!host || !out) return set_err(ctx, SPG_EINVAL, "null argument");
    *out = nullptr;
    if (host->location != SPG_HOST) return set_err(ctx, SPG_EINVAL, "upload takes a host view");
    spg_status s = check_host_csr(ctx, host, "matrix");
    if (s != SPG_OK) return s;
    SPG_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    spg_device_matrix *m = new (std::nothrow) spg_device_matrix();
    if (!m) return set_err(ctx, SPG_ENOMEM, "out of memory");
    m->device = ctx->device;
    m->rows = host->rows;
    m->cols = host->cols;
    m->nnz = host->nnz;
    const size_t rp = ((size_t)host->rows + 1) * 8;
    hipError_t e = hipMalloc(&m->row_ptr, rp);
    if (e == hipSuccess) e = hipMalloc(&m->col, host->nnz ? (size_t)host->nnz * 4 : 4);
    if (e == hipSuccess) e = hipMalloc(&m->val, host->nnz ? (size_t)host->nnz * 8 : 8);
    if (e == hipSuccess) e = hipMemcpy(m->row_ptr, host->row_ptr, rp, hipMemcpyHostToDevice);
    if (e == hipSuccess && host->nnz) e = hipMemcpy(m->col, host->col_idx, (size_t)host->nnz * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess && host->nnz) e = hipMemcpy(m->val, host->values, (size_t)host->nnz * 8, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        spg_device_matrix_free(m);
        return hip_err(ctx, e, "upload matrix");
    }
    *out = m;
    return SPG_OK;
}

spg_status spg_device_matrix_view(const spg_device_matrix *m, spg_csr_view *out) {
    if (!m || !out) return SPG_EINVAL;
    out->rows = m->rows;
    out->cols = m->cols;
    out->nnz = m->nnz;
    out->row_ptr = static_cast<const int64_t *>(m->row_ptr);
    out->col_idx = static_cast<const int32_t *>(m->col);
    out->values = static_cast<const double *>(m->val);
    out->location = SPG_DEVICE;
    return SPG_OK;
}

spg_status spg_device_matrix_free(spg_device_matrix *m) {
    if (!m) return SPG_OK;
    (void)hipSetDevice(m->device);
    if (m->row_ptr) hipFree(m->row_ptr);
    if (m->col) hipFree(m->col);
    if (m->val) hipFree(m->val);
    delete m;
    return SPG_OK;
}

spg_status spg_synchronize(spg_context *ctx) {
    if (!ctx) return SPG_EINVAL;
    SPG_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return SPG_OK;
}

}  // extern "C"
