// accumulate.hip -- staged accumulation (trident::LocalAccumulator) on the device.
//
// Reference: src/trident/common/matrix_blocks.cpp:152-200.  addProduct(a, b)
// performs acc[row][col] += a*b for A entries in stored order and, within each,
// B entries in stored order, on per-row maps that persist across calls; finish()
// drops exact zeros and returns the rows with columns ascending.
//
// Here the accumulator state is a CSR matrix that keeps every structural
// coordinate (zeros included).  One stage is one engine product whose serial
// order starts with the running state: row r of the stage's A' is the entry
// (r, 1.0) followed by A's row r with columns shifted past the state rows, and
// B' stacks the state on top of B.  A column already in the state therefore
// starts from its running value (times 1.0, which is exact) and continues with
// the stage's products left to right -- the reference's `+=` sequence.  A new
// column starts from its first product where the reference starts from
// 0.0 + product; the two differ only in the sign of a zero, and zeros are
// dropped by finish().  The stage runs with zero dropping switched off, and
// its result becomes the next state.
#include "engine_internal.hpp"

#include <algorithm>
#include <new>
#include <vector>

using namespace spg;

namespace {

// Per row: entries whose value is not 0.0.
__global__ void k_count_nonzero(int32_t rows, const int64_t *rp, const double *val, int64_t *cnt) {
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const int lane = threadIdx.x & 63;
    for (int64_t r = wave; r < rows; r += nw) {
        int64_t c = 0;
        for (int64_t q = rp[r] + lane; q < rp[r + 1]; q += 64) c += val[q] != 0.0;
        for (int d = 32; d > 0; d >>= 1) c += __shfl_xor(c, d, 64);
        if (lane == 0) cnt[r] = c;
    }
}

// Copies the nonzero entries of each row, in order, to its packed offset.
__global__ void k_drop_zeros(int32_t rows, const int64_t *rp, const int32_t *col, const double *val,
                             const int64_t *rp_out, int32_t *col_out, double *val_out) {
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const int lane = threadIdx.x & 63;
    for (int64_t r = wave; r < rows; r += nw) {
        int64_t o = rp_out[r];
        for (int64_t q0 = rp[r]; q0 < rp[r + 1]; q0 += 64) {
            const int64_t q = q0 + lane;
            const bool keep = q < rp[r + 1] && val[q] != 0.0;
            const unsigned long long m = __ballot(keep);
            if (keep) {
                const int64_t d = o + __popcll(m & ((1ull << lane) - 1ull));
                col_out[d] = col[q];
                val_out[d] = val[q];
            }
            o += __popcll(m);
        }
    }
}

int blocks_for_rows(int64_t rows) { return (int)std::max<int64_t>(1, std::min<int64_t>((rows * 64 + 255) / 256, 65536)); }

}  // namespace

struct spg_accumulator {
    spg_context *parent = nullptr;
    spg_context *own = nullptr;  // runs the stage products
    int32_t rows = 0, cols = 0;
    DevBuf s_ptr, s_col, s_val;  // state: every structural coordinate, zeros kept
    int64_t s_nnz = 0;
    DevBuf a_ptr, a_col, a_val, b_ptr, b_col, b_val;  // the stage's A' and B'
    DevBuf f_ptr, f_col, f_val, f_cnt;                // finish(): zeros dropped
    int64_t f_nnz = 0;
    bool finished = false;
    ~spg_accumulator() {
        DevBuf *bufs[] = {&s_ptr, &s_col, &s_val, &a_ptr, &a_col, &a_val, &b_ptr,
                          &b_col, &b_val, &f_ptr, &f_col, &f_val, &f_cnt};
        for (DevBuf *b : bufs) b->release();
        if (own) spg_destroy(own);
    }
};

extern "C" {

spg_status spg_accumulator_create(spg_context *ctx, int32_t rows, int32_t cols, spg_accumulator **out) {
    if (!ctx) return SPG_EINVAL;
    if (!out) return set_err(ctx, SPG_EINVAL, "null argument");
    *out = nullptr;
    if (rows < 0 || cols < 0) return set_err(ctx, SPG_EINVAL, "negative dimension");
    SPG_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    spg_accumulator *acc = new (std::nothrow) spg_accumulator();
    if (!acc) return set_err(ctx, SPG_ENOMEM, "out of memory");
    spg_status s = spg_create(&acc->own, ctx->device);
    if (s != SPG_OK) {
        delete acc;
        return set_err(ctx, s, "accumulator context");
    }
    acc->parent = ctx;
    acc->own->stream = ctx->stream;
    acc->rows = rows;
    acc->cols = cols;
    hipError_t e = acc->s_ptr.ensure(((size_t)rows + 1) * 8, ctx->stream);
    if (e != hipSuccess) {
        delete acc;
        return hip_err(ctx, e, "hipMalloc(accumulator)");
    }
    launch_zero_bytes(acc->s_ptr.p, ((size_t)rows + 1) * 8, ctx->stream);
    *out = acc;
    return SPG_OK;
}

spg_status spg_accumulator_add(spg_accumulator *acc, const spg_csr_view *A, const spg_csr_view *B) {
    if (!acc) return SPG_EINVAL;
    spg_context *ctx = acc->parent;
    if (!A || !B) return set_err(ctx, SPG_EINVAL, "null argument");
    spg_status s;
    if ((s = check_host_csr(ctx, A, "a")) != SPG_OK) return s;
    if ((s = check_host_csr(ctx, B, "b")) != SPG_OK) return s;
    if (A->location != SPG_HOST || B->location != SPG_HOST)
        return set_err(ctx, SPG_EINVAL, "staged products take host views");
    // matrix_blocks.cpp:156-158
    if (A->cols != B->rows || A->rows != acc->rows || B->cols != acc->cols)
        return set_err(ctx, SPG_EINVAL, "incompatible Trident partial product shapes");
    SPG_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    hipStream_t st = ctx->stream;
    const int64_t R = acc->rows;
    const int64_t ndir = R + B->rows;
    if (ndir > INT32_MAX) return set_err(ctx, SPG_EOVERFLOW, "accumulator rows plus B rows exceed INT32_MAX");
    // A': (r, 1.0) then A's row with columns shifted past the state rows.
    std::vector<int64_t> ap((size_t)R + 1);
    std::vector<int32_t> ac;
    std::vector<double> av;
    ac.reserve((size_t)(A->nnz + R));
    av.reserve((size_t)(A->nnz + R));
    ap[0] = 0;
    for (int64_t r = 0; r < R; ++r) {
        ac.push_back((int32_t)r);
        av.push_back(1.0);
        for (int64_t e = A->row_ptr[r]; e < A->row_ptr[r + 1]; ++e) {
            const int32_t c = A->col_idx[e];
            ac.push_back(c >= 0 && c < B->rows ? (int32_t)(c + R) : -1);  // outside B: reported as a range error
            av.push_back(A->values[e]);
        }
        ap[(size_t)r + 1] = (int64_t)ac.size();
    }
    // B': the state rows, then B's rows.
    const int64_t bn = acc->s_nnz + B->nnz;
    std::vector<int64_t> bp_tail((size_t)B->rows);
    for (int64_t r = 0; r < B->rows; ++r) bp_tail[(size_t)r] = acc->s_nnz + B->row_ptr[r + 1];
    SPG_HIP(acc->a_ptr.ensure(ap.size() * 8, st), "hipMalloc(stage)");
    SPG_HIP(acc->a_col.ensure(ac.size() * 4 + 4, st), "hipMalloc(stage)");
    SPG_HIP(acc->a_val.ensure(av.size() * 8 + 8, st), "hipMalloc(stage)");
    SPG_HIP(acc->b_ptr.ensure((size_t)(ndir + 1) * 8, st), "hipMalloc(stage)");
    SPG_HIP(acc->b_col.ensure((size_t)bn * 4 + 4, st), "hipMalloc(stage)");
    SPG_HIP(acc->b_val.ensure((size_t)bn * 8 + 8, st), "hipMalloc(stage)");
    SPG_HIP(hipMemcpyAsync(acc->a_ptr.p, ap.data(), ap.size() * 8, hipMemcpyHostToDevice, st), "upload stage");
    if (!ac.empty()) {
        SPG_HIP(hipMemcpyAsync(acc->a_col.p, ac.data(), ac.size() * 4, hipMemcpyHostToDevice, st), "upload stage");
        SPG_HIP(hipMemcpyAsync(acc->a_val.p, av.data(), av.size() * 8, hipMemcpyHostToDevice, st), "upload stage");
    }
    SPG_HIP(hipMemcpyAsync(acc->b_ptr.p, acc->s_ptr.p, (size_t)(R + 1) * 8, hipMemcpyDeviceToDevice, st), "stage");
    if (B->rows)
        SPG_HIP(hipMemcpyAsync(acc->b_ptr.as<int64_t>() + R + 1, bp_tail.data(), bp_tail.size() * 8,
                               hipMemcpyHostToDevice, st), "upload stage");
    if (acc->s_nnz) {
        SPG_HIP(hipMemcpyAsync(acc->b_col.p, acc->s_col.p, (size_t)acc->s_nnz * 4, hipMemcpyDeviceToDevice, st), "stage");
        SPG_HIP(hipMemcpyAsync(acc->b_val.p, acc->s_val.p, (size_t)acc->s_nnz * 8, hipMemcpyDeviceToDevice, st), "stage");
    }
    if (B->nnz) {
        SPG_HIP(hipMemcpyAsync(acc->b_col.as<int32_t>() + acc->s_nnz, B->col_idx, (size_t)B->nnz * 4,
                               hipMemcpyHostToDevice, st), "upload stage");
        SPG_HIP(hipMemcpyAsync(acc->b_val.as<double>() + acc->s_nnz, B->values, (size_t)B->nnz * 8,
                               hipMemcpyHostToDevice, st), "upload stage");
    }
    SPG_HIP(hipStreamSynchronize(st), "hipStreamSynchronize");  // host vectors go out of scope

    ProductPlan plan;
    Operands &op = plan.op;
    op.aptr = acc->a_ptr.as<int64_t>();
    op.acol = acc->a_col.as<int32_t>();
    op.aval = acc->a_val.as<double>();
    op.nnz_a = (int64_t)ac.size();
    op.arows = (int32_t)R;
    op.ndir = (int32_t)ndir;
    op.bstart = acc->b_ptr.as<int64_t>();
    op.bend = acc->b_ptr.as<int64_t>() + 1;
    op.bcol = acc->b_col.as<int32_t>();
    op.bval = acc->b_val.as<double>();
    op.bcols = acc->cols;
    op.keep_zeros = 1;
    plan.rows = (int32_t)R;
    plan.cols = acc->cols;
    spg_context *own = acc->own;
    own->path = ctx->path;
    own->force_wide = ctx->force_wide;
    own->large_mode = ctx->large_mode;
    own->allow_direct = false;
    plan.small_b = narrow_b(own, bn);
    plan.b_span = bn;
    if ((s = run_product(own, plan)) != SPG_OK) return set_err(ctx, s, own->err);
    // The stage's result is the new state.
    const int64_t n = own->res_nnz;
    SPG_HIP(acc->s_col.ensure((size_t)n * 4 + 4, st), "hipMalloc(accumulator)");
    SPG_HIP(acc->s_val.ensure((size_t)n * 8 + 8, st), "hipMalloc(accumulator)");
    const int c = own->cur;
    SPG_HIP(hipMemcpyAsync(acc->s_ptr.p, own->rowptr[c].p, (size_t)(R + 1) * 8, hipMemcpyDeviceToDevice, st), "state");
    if (n) {
        SPG_HIP(hipMemcpyAsync(acc->s_col.p, own->ccol[c].p, (size_t)n * 4, hipMemcpyDeviceToDevice, st), "state");
        SPG_HIP(hipMemcpyAsync(acc->s_val.p, own->cval[c].p, (size_t)n * 8, hipMemcpyDeviceToDevice, st), "state");
    }
    SPG_HIP(hipStreamSynchronize(st), "hipStreamSynchronize");
    acc->s_nnz = n;
    acc->finished = false;
    return SPG_OK;
}

spg_status spg_accumulator_finish(spg_accumulator *acc, int64_t *nnz_c) {
    if (!acc) return SPG_EINVAL;
    spg_context *ctx = acc->parent;
    if (!nnz_c) return set_err(ctx, SPG_EINVAL, "null argument");
    SPG_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    hipStream_t st = ctx->stream;
    const int64_t R = acc->rows;
    SPG_HIP(acc->f_ptr.ensure((size_t)(R + 1) * 8, st), "hipMalloc(finish)");
    SPG_HIP(acc->f_cnt.ensure((size_t)(R + 1) * 8, st), "hipMalloc(finish)");
    if (R > 0) {
        hipLaunchKernelGGL(k_count_nonzero, dim3(blocks_for_rows(R)), dim3(256), 0, st, (int32_t)R,
                           acc->s_ptr.as<int64_t>(), acc->s_val.as<double>(), acc->f_cnt.as<int64_t>());
        SPG_HIP(acc->own->scan_tmp.ensure(scan_tmp_elems(R) * 8, st), "hipMalloc(scan)");
        launch_scan(acc->f_cnt.as<int64_t>(), acc->f_ptr.as<int64_t>(), R, acc->own->scan_tmp.as<int64_t>(), nullptr,
                    st);
    } else {
        launch_zero_bytes(acc->f_ptr.p, 8, st);
    }
    int64_t n = 0;
    SPG_HIP(hipMemcpyAsync(&n, acc->f_ptr.as<int64_t>() + R, 8, hipMemcpyDeviceToHost, st), "read count");
    SPG_HIP(hipStreamSynchronize(st), "hipStreamSynchronize");
    SPG_HIP(acc->f_col.ensure((size_t)n * 4 + 4, st), "hipMalloc(finish)");
    SPG_HIP(acc->f_val.ensure((size_t)n * 8 + 8, st), "hipMalloc(finish)");
    if (R > 0 && n > 0)
        hipLaunchKernelGGL(k_drop_zeros, dim3(blocks_for_rows(R)), dim3(256), 0, st, (int32_t)R,
                           acc->s_ptr.as<int64_t>(), acc->s_col.as<int32_t>(), acc->s_val.as<double>(),
                           acc->f_ptr.as<int64_t>(), acc->f_col.as<int32_t>(), acc->f_val.as<double>());
    SPG_HIP(hipStreamSynchronize(st), "hipStreamSynchronize");
    SPG_HIP(hipGetLastError(), "finish kernels");
    acc->f_nnz = n;
    acc->finished = true;
    *nnz_c = n;
    return SPG_OK;
}

spg_status spg_accumulator_copy(spg_accumulator *acc, int64_t *row_ptr, int32_t *col_idx, double *values) {
    if (!acc) return SPG_EINVAL;
    spg_context *ctx = acc->parent;
    if (!acc->finished) return set_err(ctx, SPG_ESTATE, "finish the accumulator first");
    if (!row_ptr || (acc->f_nnz > 0 && (!col_idx || !values))) return set_err(ctx, SPG_EINVAL, "null output buffer");
    SPG_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    hipStream_t st = ctx->stream;
    SPG_HIP(hipMemcpyAsync(row_ptr, acc->f_ptr.p, ((size_t)acc->rows + 1) * 8, hipMemcpyDeviceToHost, st), "download");
    if (acc->f_nnz) {
        SPG_HIP(hipMemcpyAsync(col_idx, acc->f_col.p, (size_t)acc->f_nnz * 4, hipMemcpyDeviceToHost, st), "download");
        SPG_HIP(hipMemcpyAsync(values, acc->f_val.p, (size_t)acc->f_nnz * 8, hipMemcpyDeviceToHost, st), "download");
    }
    SPG_HIP(hipStreamSynchronize(st), "hipStreamSynchronize");
    return SPG_OK;
}

spg_status spg_accumulator_destroy(spg_accumulator *acc) {
    if (!acc) return SPG_OK;
    if (acc->parent && acc->parent->stream) (void)hipStreamSynchronize(acc->parent->stream);
    delete acc;
    return SPG_OK;
}

}  // extern "C"
