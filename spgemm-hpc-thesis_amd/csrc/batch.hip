// batch.hip -- many independent products as one device pipeline.
//
// A batch packs its products C_i = A_i * B_i into one resident arena: the B_i
// are stacked row-wise (columns unchanged) and the A_i are stacked row-wise
// with every column index shifted by the first arena row of its B_i.  The
// arena product A.B then holds every C_i as a block of consecutive rows with
// C_i's own column indices, and per-row semantics (serial summation order,
// ascending columns, exact-zero drop) are those of the standalone product, so
// each C_i is bit-identical to multiplying A_i and B_i alone.  One pipeline
// (analyze, single-pass kernel, big-row kernels) then runs over all rows of all
// products: the row bins and the single-pass tiles are the device-side work
// queue, and uneven product sizes balance across the whole chip.
//
// Once a batch has run, its plan (bins, capacities, kernel choices) is fixed
// for operands of the same structure; the pipeline can then be captured in a
// HIP graph and replayed without host synchronisation inside.  An autotuner
// times the pipeline variants (single pass with one or two rows per wave,
// two-phase, single pass over dealt 32-row tiles, single pass over packed
// tiles of small rows) on the batch and keeps the fastest.
//
// A chained batch takes its A operands from the results of another batch (the
// second product of a Galerkin triple (R.A).P): its arena rows are refreshed on
// the device from the source batch's output.
#include "engine_internal.hpp"

#include <algorithm>
#include <new>
#include <vector>

using namespace spg;

namespace {

__global__ void k_refresh(int64_t nrows, const int64_t *dst_rp, const int64_t *src_row, const int32_t *col_shift,
                          const int64_t *src_rp, const int32_t *src_col, const double *src_val, int32_t *dst_col,
                          double *dst_val, int *bad) {
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const int lane = threadIdx.x & 63;
    for (int64_t r = wave; r < nrows; r += nw) {
        const int64_t d0 = dst_rp[r], n = dst_rp[r + 1] - d0;
        const int64_t s0 = src_rp[src_row[r]];
        if (src_rp[src_row[r] + 1] - s0 != n) {  // the source's structure changed
            if (lane == 0) atomicOr(bad, 1);
            continue;
        }
        const int32_t sh = col_shift[r];
        for (int64_t q = lane; q < n; q += 64) {
            dst_col[d0 + q] = src_col[s0 + q] + sh;
            dst_val[d0 + q] = src_val[s0 + q];
        }
    }
}

// One chunk of arena packing: entries [e0, e1) of a product (columns shifted
// and checked, values copied), or row offsets [e0, e1) of a product (shifted
// by the product's first arena entry).
struct PackChunk {
    const int64_t *src_rp;
    const int32_t *src_col;
    const double *src_val;
    int64_t e0, e1;
    int64_t dst_off;  // arena entry (entries) or arena row (row offsets) of the product's first
    int64_t add;      // row offsets: the product's first arena entry
    int32_t shift;    // entries: column shift
    int32_t lim;      // entries: columns must lie in [0, lim) or become -1 (outside the directory)
    int32_t rows;     // 1: row-offset chunk
    int32_t pad;
};

__global__ void k_pack(const PackChunk *chunks, int64_t *dst_rp, int32_t *dst_col, double *dst_val) {
    const PackChunk c = chunks[blockIdx.x];
    for (int64_t q = c.e0 + threadIdx.x; q < c.e1; q += blockDim.x) {
        if (c.rows) {
            dst_rp[c.dst_off + q] = c.src_rp[q] + c.add;
        } else {
            const int32_t v = c.src_col[q];
            dst_col[c.dst_off + q] = c.lim == INT_MAX ? v : ((v >= 0 && v < c.lim) ? v + c.shift : -1);
            dst_val[c.dst_off + q] = c.src_val[q];
        }
    }
}

constexpr int64_t kPackChunk = 1 << 16;

}  // namespace

struct spg_batch {
    spg_context *parent = nullptr;
    spg_context *own = nullptr;  // private result buffers and workspace, parent's stream
    int32_t n = 0;
    std::vector<int64_t> a_row0;  // first arena row of each product (rows of C)
    std::vector<int32_t> rows, cols;
    int64_t total_rows = 0, a_nnz = 0, b_rows = 0, b_nnz = 0;
    int32_t max_cols = 0;
    DevBuf a_ptr, a_col, a_val, b_ptr, b_col, b_val;
    // chained batches: source rows and column shifts per arena row
    const spg_batch *src = nullptr;
    DevBuf src_row, col_shift, bad;
    ProductPlan plan;
    bool planned = false;
    int counted = 0;  // 1: the count phase ran (counts in `counts`); 2: it finished the product (direct pass)
    Control counts{};
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~spg_batch() {
        drop_graph();
        DevBuf *bufs[] = {&a_ptr, &a_col, &a_val, &b_ptr, &b_col, &b_val, &src_row, &col_shift, &bad};
        for (DevBuf *b : bufs) b->release();
        if (own) spg_destroy(own);
    }
    void drop_graph() {
        if (exec) hipGraphExecDestroy(exec);
        if (graph) hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
};

namespace {

spg_status upload(spg_context *ctx, DevBuf &buf, const void *src, size_t bytes) {
    SPG_HIP(buf.ensure(bytes ? bytes : 8, ctx->stream), "hipMalloc(batch arena)");
    if (bytes) SPG_HIP(hipMemcpyAsync(buf.p, src, bytes, hipMemcpyHostToDevice, ctx->stream), "upload batch arena");
    return SPG_OK;
}

spg_status new_batch(spg_context *ctx, spg_batch **out) {
    spg_batch *b = new (std::nothrow) spg_batch();
    if (!b) return set_err(ctx, SPG_ENOMEM, "out of memory");
    spg_status s = spg_create(&b->own, ctx->device);
    if (s != SPG_OK) {
        delete b;
        return set_err(ctx, s, "batch context");
    }
    b->parent = ctx;
    b->own->for_each_buffer([](DevBuf &d) { d.headroom = true; });
    DevBuf *arena[] = {&b->a_ptr, &b->a_col, &b->a_val, &b->b_ptr, &b->b_col, &b->b_val};
    for (DevBuf *d : arena) d->headroom = true;
    b->own->stream = ctx->stream;
    b->own->path = ctx->path;
    b->own->fused_rows = ctx->fused_rows;
    b->own->force_wide = ctx->force_wide;
    b->own->large_mode = ctx->large_mode;
    b->own->allow_direct = false;  // batch plans are replayed as analysis + values
    b->own->timing = ctx->timing;
    *out = b;
    return SPG_OK;
}

void set_operands(spg_batch *b) {
    Operands op{};
    op.aptr = b->a_ptr.as<int64_t>();
    op.acol = b->a_col.as<int32_t>();
    op.aval = b->a_val.as<double>();
    op.nnz_a = b->a_nnz;
    op.arows = (int32_t)b->total_rows;
    op.ndir = (int32_t)b->b_rows;
    op.bstart = b->b_ptr.as<int64_t>();
    op.bend = b->b_ptr.as<int64_t>() + 1;
    op.bcol = b->b_col.as<int32_t>();
    op.bval = b->b_val.as<double>();
    op.bcols = b->max_cols;
    b->plan.op = op;
    b->plan.rows = (int32_t)b->total_rows;
    b->plan.cols = b->max_cols;
    b->plan.small_b = narrow_b(b->own, b->b_nnz);
    b->plan.b_span = b->b_nnz;
}

spg_status stack_b(spg_context *ctx, spg_batch *b, int32_t n, const spg_csr_view *B, std::vector<int64_t> &b_row0) {
    std::vector<int64_t> bp(1, 0);
    std::vector<int32_t> bc;
    std::vector<double> bv;
    b_row0.resize((size_t)n);
    for (int32_t i = 0; i < n; ++i) {
        const spg_csr_view &m = B[i];
        if (m.location != SPG_HOST) return set_err(ctx, SPG_EINVAL, "batch operands take host views");
        spg_status s = check_host_csr(ctx, &m, "B_i");
        if (s != SPG_OK) return s;
        b_row0[(size_t)i] = (int64_t)bp.size() - 1;
        const int64_t base = bp.back();
        for (int32_t r = 0; r < m.rows; ++r) bp.push_back(base + m.row_ptr[r + 1]);
        bc.insert(bc.end(), m.col_idx, m.col_idx + m.nnz);
        bv.insert(bv.end(), m.values, m.values + m.nnz);
        b->max_cols = std::max(b->max_cols, m.cols);
        b->cols.push_back(m.cols);
    }
    if ((int64_t)bp.size() - 1 > INT32_MAX) return set_err(ctx, SPG_EOVERFLOW, "batch B rows exceed INT32_MAX");
    b->b_rows = (int64_t)bp.size() - 1;
    b->b_nnz = bp.back();
    spg_status s;
    if ((s = upload(ctx, b->b_ptr, bp.data(), bp.size() * 8)) != SPG_OK) return s;
    if ((s = upload(ctx, b->b_col, bc.data(), bc.size() * 4)) != SPG_OK) return s;
    if ((s = upload(ctx, b->b_val, bv.data(), bv.size() * 8)) != SPG_OK) return s;
    return SPG_OK;
}

// Packs n device views into an arena (row offsets, entries) with the given
// per-product first arena row/entry, column shift and column limit.
spg_status pack_device(spg_context *ctx, int32_t n, const spg_csr_view *M, const std::vector<int64_t> &row0,
                       const std::vector<int64_t> &ent0, const std::vector<int32_t> &shift,
                       const std::vector<int32_t> &lim, DevBuf &rp, DevBuf &col, DevBuf &val, DevBuf &desc) {
    std::vector<PackChunk> ch;
    for (int32_t i = 0; i < n; ++i) {
        const spg_csr_view &m = M[i];
        for (int64_t e = 0; e < m.nnz; e += kPackChunk) {
            PackChunk c{};
            c.src_rp = m.row_ptr;
            c.src_col = m.col_idx;
            c.src_val = m.values;
            c.e0 = e;
            c.e1 = std::min<int64_t>(m.nnz, e + kPackChunk);
            c.dst_off = ent0[(size_t)i];
            c.shift = shift[(size_t)i];
            c.lim = lim[(size_t)i];
            ch.push_back(c);
        }
        // row offsets 1..rows (offset 0 is the previous product's last)
        for (int64_t r = (i == 0 ? 0 : 1); r <= m.rows; r += kPackChunk) {
            PackChunk c{};
            c.src_rp = m.row_ptr;
            c.e0 = r;
            c.e1 = std::min<int64_t>((int64_t)m.rows + 1, r + kPackChunk);
            c.dst_off = row0[(size_t)i];
            c.add = ent0[(size_t)i];
            c.rows = 1;
            ch.push_back(c);
        }
    }
    if (ch.empty()) return SPG_OK;
    SPG_HIP(desc.ensure(ch.size() * sizeof(PackChunk), ctx->stream), "hipMalloc(pack descriptors)");
    SPG_HIP(hipMemcpyAsync(desc.p, ch.data(), ch.size() * sizeof(PackChunk), hipMemcpyHostToDevice, ctx->stream),
            "upload pack descriptors");
    hipLaunchKernelGGL(k_pack, dim3((unsigned)ch.size()), dim3(256), 0, ctx->stream, desc.as<PackChunk>(),
                       rp.as<int64_t>(), col.as<int32_t>(), val.as<double>());
    SPG_HIP(hipGetLastError(), "pack kernel");
    return SPG_OK;
}

// Arena from device views: the caller guarantees they stay valid during the call.
spg_status create_from_device(spg_context *ctx, spg_batch *b, int32_t n, const spg_csr_view *A,
                              const spg_csr_view *B) {
    std::vector<int64_t> arow0, aent0, brow0, bent0;
    std::vector<int32_t> ashift, alim, bshift, blim;
    int64_t ar = 0, ae = 0, br = 0, be = 0;
    for (int32_t i = 0; i < n; ++i) {
        if (A[i].location != SPG_DEVICE || B[i].location != SPG_DEVICE)
            return set_err(ctx, SPG_EINVAL, "batch operands must be all host or all device views");
        spg_status s = check_host_csr(ctx, &A[i], "A_i");
        if (s == SPG_OK) s = check_host_csr(ctx, &B[i], "B_i");
        if (s != SPG_OK) return s;
        if (A[i].cols != B[i].rows)
            return set_err(ctx, SPG_EINVAL, "SpGEMM requires A.cols == B.rows (product " + std::to_string(i) + ")");
        b->a_row0.push_back(ar);
        b->rows.push_back(A[i].rows);
        b->cols.push_back(B[i].cols);
        b->max_cols = std::max(b->max_cols, B[i].cols);
        arow0.push_back(ar);
        aent0.push_back(ae);
        brow0.push_back(br);
        bent0.push_back(be);
        ashift.push_back((int32_t)br);
        alim.push_back(A[i].cols);
        bshift.push_back(0);
        blim.push_back(INT_MAX);
        ar += A[i].rows;
        ae += A[i].nnz;
        br += B[i].rows;
        be += B[i].nnz;
    }
    if (ar > INT32_MAX || br > INT32_MAX) return set_err(ctx, SPG_EOVERFLOW, "batch rows exceed INT32_MAX");
    b->total_rows = ar;
    b->a_nnz = ae;
    b->b_rows = br;
    b->b_nnz = be;
    hipStream_t st = ctx->stream;
    SPG_HIP(b->a_ptr.ensure((size_t)(ar + 1) * 8, st), "hipMalloc(batch arena)");
    SPG_HIP(b->a_col.ensure((size_t)(ae ? ae : 1) * 4, st), "hipMalloc(batch arena)");
    SPG_HIP(b->a_val.ensure((size_t)(ae ? ae : 1) * 8, st), "hipMalloc(batch arena)");
    SPG_HIP(b->b_ptr.ensure((size_t)(br + 1) * 8, st), "hipMalloc(batch arena)");
    SPG_HIP(b->b_col.ensure((size_t)(be ? be : 1) * 4, st), "hipMalloc(batch arena)");
    SPG_HIP(b->b_val.ensure((size_t)(be ? be : 1) * 8, st), "hipMalloc(batch arena)");
    launch_zero_bytes(b->a_ptr.p, 8, st);
    launch_zero_bytes(b->b_ptr.p, 8, st);
    spg_status s;
    DevBuf desc_a, desc_b;
    if ((s = pack_device(ctx, n, A, arow0, aent0, ashift, alim, b->a_ptr, b->a_col, b->a_val, desc_a)) == SPG_OK)
        s = pack_device(ctx, n, B, brow0, bent0, bshift, blim, b->b_ptr, b->b_col, b->b_val, desc_b);
    hipError_t e = hipStreamSynchronize(st);
    desc_a.release();
    desc_b.release();
    if (s != SPG_OK) return s;
    if (e != hipSuccess) return hip_err(ctx, e, "pack batch arena");
    return SPG_OK;
}

spg_status capture(spg_batch *b) {
    spg_context *ctx = b->own;
    b->drop_graph();
    SPG_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
    enqueue_counts(ctx, b->plan, b->plan.fused ? kBigBins : kAllBins, !b->plan.fused);
    enqueue_values(ctx, b->plan);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (e != hipSuccess) return hip_err(ctx, e, "hipStreamEndCapture");
    b->graph = g;
    SPG_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0), "hipGraphInstantiate");
    return SPG_OK;
}

}  // namespace

extern "C" {

namespace {

// (Re)builds a batch's arena from n products; the batch keeps its buffers and
// private context, so a rebuilt batch allocates nothing once warm.
spg_status build_arena(spg_context *ctx, spg_batch *b, int32_t n, const spg_csr_view *A, const spg_csr_view *B) {
    b->drop_graph();
    b->planned = false;
    b->counted = 0;
    b->plan = ProductPlan{};
    b->n = n;
    b->a_row0.clear();
    b->rows.clear();
    b->cols.clear();
    b->max_cols = 0;
    b->own->have_result = false;
    spg_status s;
    if (n > 0 && A[0].location == SPG_DEVICE) {
        if ((s = create_from_device(ctx, b, n, A, B)) != SPG_OK) return s;
        set_operands(b);
        return SPG_OK;
    }
    std::vector<int64_t> b_row0;
    if ((s = stack_b(ctx, b, n, B, b_row0)) != SPG_OK) return s;
    std::vector<int64_t> ap(1, 0);
    std::vector<int32_t> ac;
    std::vector<double> av;
    for (int32_t i = 0; i < n; ++i) {
        const spg_csr_view &m = A[i];
        if (m.location != SPG_HOST) return set_err(ctx, SPG_EINVAL, "batch operands must be all host or all device views");
        if ((s = check_host_csr(ctx, &m, "A_i")) != SPG_OK) return s;
        if (m.cols != B[i].rows)
            return set_err(ctx, SPG_EINVAL, "SpGEMM requires A.cols == B.rows (product " + std::to_string(i) + ")");
        b->a_row0.push_back((int64_t)ap.size() - 1);
        b->rows.push_back(m.rows);
        const int64_t base = ap.back();
        for (int32_t r = 0; r < m.rows; ++r) ap.push_back(base + m.row_ptr[r + 1]);
        const int32_t shift = (int32_t)b_row0[(size_t)i];
        for (int64_t e = 0; e < m.nnz; ++e) {
            const int32_t c = m.col_idx[e];
            // columns outside B_i are moved outside the arena directory, so
            // the analyze pass reports them as it does for a single product
            ac.push_back(c >= 0 && c < m.cols ? c + shift : -1);
        }
        av.insert(av.end(), m.values, m.values + m.nnz);
    }
    if ((int64_t)ap.size() - 1 > INT32_MAX) return set_err(ctx, SPG_EOVERFLOW, "batch rows exceed INT32_MAX");
    b->total_rows = (int64_t)ap.size() - 1;
    b->a_nnz = ap.back();
    if ((s = upload(ctx, b->a_ptr, ap.data(), ap.size() * 8)) != SPG_OK ||
        (s = upload(ctx, b->a_col, ac.data(), ac.size() * 4)) != SPG_OK ||
        (s = upload(ctx, b->a_val, av.data(), av.size() * 8)) != SPG_OK)
        return s;
    const hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_err(ctx, e, "upload batch");
    set_operands(b);
    return SPG_OK;
}

}  // namespace

spg_status spg_batch_create(spg_context *ctx, int32_t n, const spg_csr_view *A, const spg_csr_view *B,
                            spg_batch **out) {
    if (!ctx) return SPG_EINVAL;
    if (!out || n < 0 || (n > 0 && (!A || !B))) return set_err(ctx, SPG_EINVAL, "null argument");
    *out = nullptr;
    SPG_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    spg_batch *b = nullptr;
    spg_status s = new_batch(ctx, &b);
    if (s != SPG_OK) return s;
    if ((s = build_arena(ctx, b, n, A, B)) != SPG_OK) {
        delete b;
        return s;
    }
    *out = b;
    return SPG_OK;
}

spg_status spg_batch_rebuild(spg_batch *b, int32_t n, const spg_csr_view *A, const spg_csr_view *B) {
    if (!b) return SPG_EINVAL;
    spg_context *ctx = b->parent;
    if (b->src) return set_err(ctx, SPG_ESTATE, "a chained batch cannot be rebuilt");
    if (n < 0 || (n > 0 && (!A || !B))) return set_err(ctx, SPG_EINVAL, "null argument");
    SPG_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    SPG_HIP(hipStreamSynchronize(b->own->stream), "hipStreamSynchronize");
    return build_arena(ctx, b, n, A, B);
}

spg_status spg_batch_create_chained(spg_context *ctx, const spg_batch *src, int32_t n, const int32_t *src_index,
                                    const spg_csr_view *B, spg_batch **out) {
    if (!ctx) return SPG_EINVAL;
    if (!src || !out || n < 0 || (n > 0 && (!src_index || !B))) return set_err(ctx, SPG_EINVAL, "null argument");
    if (!src->own->have_result) return set_err(ctx, SPG_ESTATE, "the source batch has not run");
    *out = nullptr;
    spg_batch *b = nullptr;
    spg_status s = new_batch(ctx, &b);
    if (s != SPG_OK) return s;
    b->n = n;
    b->src = src;
    std::vector<int64_t> b_row0;
    if ((s = stack_b(ctx, b, n, B, b_row0)) != SPG_OK) {
        delete b;
        return s;
    }
    // The source's row offsets give each chained A_i's structure.
    const spg_context *sc = src->own;
    std::vector<int64_t> srp((size_t)sc->res_rows + 1);
    SPG_HIP(hipMemcpy(srp.data(), sc->rowptr[sc->cur].p, srp.size() * 8, hipMemcpyDeviceToHost), "download rows");
    std::vector<int64_t> ap(1, 0), srow;
    std::vector<int32_t> shift;
    for (int32_t i = 0; i < n; ++i) {
        const int32_t k = src_index[i];
        if (k < 0 || k >= src->n) {
            delete b;
            return set_err(ctx, SPG_EINVAL, "source product index out of range");
        }
        if (src->cols[(size_t)k] != B[i].rows) {
            delete b;
            return set_err(ctx, SPG_EINVAL, "SpGEMM requires A.cols == B.rows (product " + std::to_string(i) + ")");
        }
        b->a_row0.push_back((int64_t)ap.size() - 1);
        b->rows.push_back(src->rows[(size_t)k]);
        for (int32_t r = 0; r < src->rows[(size_t)k]; ++r) {
            const int64_t g = src->a_row0[(size_t)k] + r;
            ap.push_back(ap.back() + (srp[(size_t)g + 1] - srp[(size_t)g]));
            srow.push_back(g);
            shift.push_back((int32_t)b_row0[(size_t)i]);
        }
    }
    b->total_rows = (int64_t)ap.size() - 1;
    b->a_nnz = ap.back();
    if ((s = upload(ctx, b->a_ptr, ap.data(), ap.size() * 8)) != SPG_OK ||
        (s = upload(ctx, b->src_row, srow.data(), srow.size() * 8)) != SPG_OK ||
        (s = upload(ctx, b->col_shift, shift.data(), shift.size() * 4)) != SPG_OK) {
        delete b;
        return s;
    }
    SPG_HIP(b->a_col.ensure((size_t)(b->a_nnz ? b->a_nnz : 1) * 4, ctx->stream), "hipMalloc(batch arena)");
    SPG_HIP(b->a_val.ensure((size_t)(b->a_nnz ? b->a_nnz : 1) * 8, ctx->stream), "hipMalloc(batch arena)");
    SPG_HIP(b->bad.ensure(8, ctx->stream), "hipMalloc(flag)");
    set_operands(b);
    *out = b;
    s = spg_batch_refresh(b);
    if (s != SPG_OK) {
        delete b;
        *out = nullptr;
    }
    return s;
}

spg_status spg_batch_refresh(spg_batch *b) {
    if (!b) return SPG_EINVAL;
    spg_context *ctx = b->parent;
    if (!b->src) return set_err(ctx, SPG_ESTATE, "not a chained batch");
    const spg_context *sc = b->src->own;
    if (!sc->have_result) return set_err(ctx, SPG_ESTATE, "the source batch has not run");
    b->counted = 0;  // A's columns change: counts from before are void
    launch_zero_bytes(b->bad.p, 8, ctx->stream);
    if (b->total_rows > 0) {
        const int64_t blocks = std::min<int64_t>((b->total_rows * 64 + 255) / 256, 65536);
        hipLaunchKernelGGL(k_refresh, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, b->total_rows,
                           b->a_ptr.as<int64_t>(), b->src_row.as<int64_t>(), b->col_shift.as<int32_t>(),
                           sc->rowptr[sc->cur].as<int64_t>(), sc->ccol[sc->cur].as<int32_t>(),
                           sc->cval[sc->cur].as<double>(), b->a_col.as<int32_t>(), b->a_val.as<double>(),
                           b->bad.as<int>());
    }
    int bad = 0;
    SPG_HIP(hipMemcpyAsync(&bad, b->bad.p, 4, hipMemcpyDeviceToHost, ctx->stream), "read flag");
    SPG_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    if (bad) return set_err(ctx, SPG_ESTATE, "the source batch's result structure changed; create the chained batch again");
    return SPG_OK;
}

spg_status spg_batch_run(spg_batch *b, int64_t *nnz_total) {
    if (!b || !nnz_total) return SPG_EINVAL;
    spg_context *ctx = b->own;
    spg_status s;
    if (b->counted) {
        const int counted = b->counted;
        b->counted = 0;
        if (counted == 1 && (s = value_product(ctx, b->plan, b->counts)) != SPG_OK) return s;
        b->planned = true;
    } else if (!b->planned || !b->use_graph) {
        if ((s = run_product(ctx, b->plan)) != SPG_OK) return s;
        b->planned = true;
    } else {
        if (!b->exec && (s = capture(b)) != SPG_OK) return s;
        ctx->have_result = false;
        if (ctx->timing) SPG_HIP(hipEventRecord(ctx->ev[0], ctx->stream), "hipEventRecord");
        SPG_HIP(hipGraphLaunch(b->exec, ctx->stream), "hipGraphLaunch");
        if (ctx->timing) SPG_HIP(hipEventRecord(ctx->ev[3], ctx->stream), "hipEventRecord");
        if ((s = read_control(ctx)) != SPG_OK) return s;
        spg_stats stats{};
        if ((s = finish_product(ctx, b->plan, *ctx->h_ctrl, stats)) != SPG_OK) return s;
        if (ctx->timing) {
            float t = 0;
            SPG_HIP(hipEventElapsedTime(&t, ctx->ev[0], ctx->ev[3]), "hipEventElapsedTime");
            ctx->stats.ms_total = ctx->stats.ms_numeric = t;
        }
    }
    *nnz_total = ctx->res_nnz;
    return SPG_OK;
}

spg_status spg_batch_count(spg_batch *b, uint64_t *value_bytes_out) {
    if (!b || !value_bytes_out) return SPG_EINVAL;
    spg_context *ctx = b->own;
    b->counted = 0;
    bool done = false;
    spg_status s = count_product(ctx, b->plan, b->counts, done);
    if (s != SPG_OK) return s;
    b->counted = done ? 2 : 1;
    *value_bytes_out = done ? 0 : value_bytes(ctx, b->plan, b->counts);
    return SPG_OK;
}

spg_status spg_batch_device_bytes(const spg_batch *b, uint64_t *bytes) {
    if (!b || !bytes) return SPG_EINVAL;
    uint64_t n = b->own->device_bytes();
    const DevBuf *bufs[] = {&b->a_ptr, &b->a_col, &b->a_val, &b->b_ptr, &b->b_col, &b->b_val,
                            &b->src_row, &b->col_shift, &b->bad};
    for (const DevBuf *d : bufs) n += d->bytes;
    *bytes = n;
    return SPG_OK;
}

spg_status spg_batch_estimate(spg_context *ctx, const spg_batch *reuse, int64_t rows, int64_t a_nnz, int64_t b_rows,
                              int64_t b_nnz, int32_t cols, int64_t c_entries, uint64_t *bytes) {
    if (!ctx) return SPG_EINVAL;
    if (!bytes || rows < 0 || rows > INT32_MAX || a_nnz < 0 || b_rows < 0 || b_rows > INT32_MAX || b_nnz < 0 ||
        cols < 0 || c_entries < 0)
        return set_err(ctx, SPG_EINVAL, "invalid batch sizes");
    spg_batch *b = const_cast<spg_batch *>(reuse);  // read only: the buffers' current sizes
    spg_context fresh;                               // a new batch's (empty) buffers
    fresh.total_mem = ctx->total_mem;
    fresh.cus = ctx->cus;
    spg_context *own = b ? b->own : &fresh;
    DevBuf none[6];
    DevBuf *arena[6] = {&none[0], &none[1], &none[2], &none[3], &none[4], &none[5]};
    if (b) {
        DevBuf *mine[6] = {&b->a_ptr, &b->a_col, &b->a_val, &b->b_ptr, &b->b_col, &b->b_val};
        std::copy(mine, mine + 6, arena);
    }
    BufferNeeds v;
    v.add(*arena[0], (size_t)(rows + 1) * 8, "");
    v.add(*arena[1], (size_t)(a_nnz ? a_nnz : 1) * 4, "");
    v.add(*arena[2], (size_t)(a_nnz ? a_nnz : 1) * 8, "");
    v.add(*arena[3], (size_t)(b_rows + 1) * 8, "");
    v.add(*arena[4], (size_t)(b_nnz ? b_nnz : 1) * 4, "");
    v.add(*arena[5], (size_t)(b_nnz ? b_nnz : 1) * 8, "");
    int64_t lwidth = 0;
    int lblocks = 0;
    workspace_needs(own, (int32_t)rows, a_nnz, (int32_t)b_rows, b_nnz < (int64_t(1) << 32) && b_rows > 0, cols, v,
                    lwidth, lblocks);
    if (c_entries > 0) {
        v.add(own->ccol[0], (size_t)c_entries * 4, "");
        v.add(own->cval[0], (size_t)c_entries * 8, "");
    }
    *bytes = v.growth();
    return SPG_OK;
}

spg_status spg_batch_trim(spg_batch *b) {
    if (!b) return SPG_EINVAL;
    if (b->src) return set_err(b->own, SPG_ESTATE, "a chained batch cannot be trimmed");
    spg_context *ctx = b->own;
    SPG_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    SPG_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    b->drop_graph();
    DevBuf *bufs[] = {&b->a_ptr, &b->a_col, &b->a_val, &b->b_ptr, &b->b_col, &b->b_val};
    for (DevBuf *d : bufs) d->release();
    ctx->for_each_buffer([](DevBuf &d) { d.release(); });
    ctx->lbits_width = 0;
    ctx->have_result = false;
    b->n = 0;
    b->a_row0.clear();
    b->rows.clear();
    b->cols.clear();
    b->total_rows = b->a_nnz = b->b_rows = b->b_nnz = 0;
    b->planned = false;
    b->counted = 0;
    b->plan = ProductPlan{};
    return SPG_OK;
}

spg_status spg_batch_set_stream(spg_batch *b, void *stream) {
    if (!b) return SPG_EINVAL;
    b->own->stream = stream ? static_cast<hipStream_t>(stream) : b->parent->stream;
    return SPG_OK;
}

const char *spg_batch_last_error(const spg_batch *b) { return b ? b->own->err.c_str() : "null batch"; }

spg_status spg_batch_set_graph(spg_batch *b, int enabled) {
    if (!b) return SPG_EINVAL;
    b->use_graph = enabled != 0;
    if (!b->use_graph) b->drop_graph();
    return SPG_OK;
}

spg_status spg_batch_autotune(spg_batch *b, int repeats, int *choice) {
    if (!b) return SPG_EINVAL;
    spg_context *ctx = b->own;
    if (repeats < 1) repeats = 1;
    struct Variant {
        int path, rows;
    };
    const Variant variants[] = {{SPG_PATH_SINGLE_PASS, 2},
                                {SPG_PATH_SINGLE_PASS, 1},
                                {SPG_PATH_TWO_PHASE, 2},
                                {SPG_PATH_SINGLE_PASS, kFusedDealRows},
                                {SPG_PATH_SINGLE_PASS, kFusedPackRows}};
    constexpr int kVariants = (int)(sizeof(variants) / sizeof(variants[0]));
    const bool was_graph = b->use_graph;
    b->counted = 0;
    b->use_graph = false;
    b->drop_graph();
    float best = 0;
    int best_i = 0;
    for (int i = 0; i < kVariants; ++i) {
        ctx->path = variants[i].path;
        ctx->fused_rows = variants[i].rows;
        int64_t nnz = 0;
        spg_status s = spg_batch_run(b, &nnz);  // plans the variant
        if (s != SPG_OK) return s;
        hipEvent_t e0, e1;
        SPG_HIP(hipEventCreate(&e0), "hipEventCreate");
        SPG_HIP(hipEventCreate(&e1), "hipEventCreate");
        float total = 0;
        for (int r = 0; r < repeats; ++r) {
            SPG_HIP(hipEventRecord(e0, ctx->stream), "hipEventRecord");
            enqueue_counts(ctx, b->plan, b->plan.fused ? kBigBins : kAllBins, !b->plan.fused);
            enqueue_values(ctx, b->plan);
            SPG_HIP(hipEventRecord(e1, ctx->stream), "hipEventRecord");
            SPG_HIP(hipEventSynchronize(e1), "hipEventSynchronize");
            float t = 0;
            SPG_HIP(hipEventElapsedTime(&t, e0, e1), "hipEventElapsedTime");
            total += t;
        }
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        if (i == 0 || total < best) {
            best = total;
            best_i = i;
        }
    }
    ctx->path = variants[best_i].path;
    ctx->fused_rows = variants[best_i].rows;
    int64_t nnz = 0;
    spg_status s = spg_batch_run(b, &nnz);  // re-plan with the chosen variant
    b->use_graph = was_graph;
    if (choice) *choice = best_i;
    return s;
}

spg_status spg_batch_product_nnz(const spg_batch *b, int32_t i, int64_t *nnz) {
    if (!b || !nnz) return SPG_EINVAL;
    const spg_context *ctx = b->own;
    if (!ctx->have_result) return SPG_ESTATE;
    if (i < 0 || i >= b->n) return SPG_EINVAL;
    int64_t v[2] = {0, 0};
    const int64_t r0 = b->a_row0[(size_t)i];
    hipError_t e = hipMemcpy(v, ctx->rowptr[ctx->cur].as<int64_t>() + r0, 8, hipMemcpyDeviceToHost);
    if (e == hipSuccess)
        e = hipMemcpy(v + 1, ctx->rowptr[ctx->cur].as<int64_t>() + r0 + b->rows[(size_t)i], 8, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return SPG_EDEVICE;
    }
    *nnz = v[1] - v[0];
    return SPG_OK;
}

spg_status spg_batch_copy(spg_batch *b, int32_t i, int64_t *row_ptr, int32_t *col_idx, double *values) {
    if (!b || !row_ptr) return SPG_EINVAL;
    if (i < 0 || i >= b->n) return set_err(b->own, SPG_EINVAL, "product index out of range");
    const int64_t r0 = b->a_row0[(size_t)i];
    int64_t cnt = 0;
    return spg_result_rows(b->own, (int32_t)r0, (int32_t)(r0 + b->rows[(size_t)i]), row_ptr, col_idx, values,
                           INT64_MAX, &cnt);
}

spg_status spg_batch_info(const spg_batch *b, int32_t i, int32_t *rows, int32_t *cols) {
    if (!b || i < 0 || i >= b->n) return SPG_EINVAL;
    if (rows) *rows = b->rows[(size_t)i];
    if (cols) *cols = b->cols[(size_t)i];
    return SPG_OK;
}

spg_status spg_batch_last_stats(const spg_batch *b, spg_stats *out) {
    if (!b || !out) return SPG_EINVAL;
    *out = b->own->stats;
    return SPG_OK;
}

spg_status spg_batch_destroy(spg_batch *b) {
    if (!b) return SPG_OK;
    if (b->own && b->own->stream) (void)hipStreamSynchronize(b->own->stream);
    delete b;
    return SPG_OK;
}

}  // extern "C"
