// halo.hip -- device-resident halo rows: the global-row directory of a
// product whose B rows are a local block plus received halo slots, and the
// packing of requested local rows for sending.
//
// Receiver side, bRowForColumn (src/baselines/spgemm_exchange.cpp:75-87):
// global B row k is local row k - first_row inside the local block, else the
// halo slot that holds k.  The directory gives every k its entry range in one
// array [B_local entries | halo entries]; rows in neither stay at -1, which
// the analysis reports as "missing a remote sparse row required by A".  Built
// on the device from device views, so halo rows received straight into HBM
// (RCCL, IPC) never cross PCIe.
//
// Sender side (prepareRequestedRows and the packing loop of
// exchangeRemoteSparseRowsTwoSided, spgemm_exchange.cpp:60-73, 217-235): the
// requested global rows, each checked to be local, become one contiguous CSR
// of their entries in request order -- the payload the two-sided exchange
// sends -- with per-row counts and offsets.
#include "device_common.hpp"
#include "engine_internal.hpp"

namespace spg {
namespace {

using namespace dev;

constexpr unsigned kHaloRange = 1u;       // a halo slot's global row is outside A's columns
constexpr unsigned kHaloDuplicate = 2u;   // two halo slots hold the same global row
constexpr unsigned kHaloBadPtr = 4u;      // halo row_ptr negative or decreasing
constexpr unsigned kPackNotLocal = 8u;    // a requested row is not in the local block

__global__ void k_fill_i64(int64_t *p, int64_t n, int64_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_zero_u32(unsigned *p) { *p = 0u; }

__global__ void k_dir_local(const int64_t *bptr, int32_t rows, int32_t first, int64_t *ds, int64_t *de) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += (int64_t)gridDim.x * blockDim.x) {
        ds[first + r] = bptr[r];
        de[first + r] = bptr[r + 1];
    }
}

// Local rows take precedence over a halo slot for the same row (as on the
// host path); a row in two halo slots is an error.
__global__ void k_dir_halo(const int32_t *grows, const int64_t *hptr, int32_t n, int32_t first, int32_t local_rows,
                           int32_t ndir, int64_t base, int64_t *ds, int64_t *de, unsigned *err) {
    unsigned flags = 0;
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < n; s += (int64_t)gridDim.x * blockDim.x) {
        const int32_t g = grows[s];
        const int64_t h0 = hptr[s], h1 = hptr[s + 1];
        if (h1 < h0 || h0 < 0) flags |= kHaloBadPtr;
        if (g < 0 || g >= ndir) {
            flags |= kHaloRange;
            continue;
        }
        if (g >= first && g < first + local_rows) continue;
        const unsigned long long old =
            atomicCAS(reinterpret_cast<unsigned long long *>(&ds[g]), (unsigned long long)(int64_t)-1,
                      (unsigned long long)(base + h0));
        if (old != (unsigned long long)(int64_t)-1) {
            flags |= kHaloDuplicate;
            continue;
        }
        de[g] = base + h1;
    }
    if (flags) atomicOr(err, flags);
}

__global__ void k_pack_counts(const int64_t *bptr, int32_t rows, int32_t first, const int32_t *req, int32_t n,
                              int64_t *cnt, unsigned *err) {
    unsigned flags = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t local = (int64_t)req[i] - first;
        if (local < 0 || local >= rows) {
            flags |= kPackNotLocal;
            cnt[i] = 0;
        } else {
            cnt[i] = bptr[local + 1] - bptr[local];
        }
    }
    if (flags) atomicOr(err, flags);
}

// One wave per requested row: its entries, coalesced, to their packed place.
__global__ __launch_bounds__(256) void k_pack_rows(const int64_t *bptr, const int32_t *bcol, const double *bval,
                                                   int32_t first, const int32_t *req, int32_t n, const int64_t *optr,
                                                   int32_t *ocol, double *oval) {
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t i = wave; i < n; i += nw) {
        const int64_t src = bptr[(int64_t)req[i] - first];
        const int64_t dst = optr[i];
        const int64_t len = optr[i + 1] - dst;
        for (int64_t q = lane_id(); q < len; q += 64) {
            ocol[dst + q] = bcol[src + q];
            oval[dst + q] = bval[src + q];
        }
    }
}

unsigned grid_for(int64_t n, int per) {
    int64_t g = (n + per - 1) / per;
    if (g < 1) g = 1;
    if (g > 65536) g = 65536;
    return (unsigned)g;
}

spg_status read_flags(spg_context *ctx, unsigned &flags) {
    SPG_HIP(hipMemcpyAsync(&flags, ctx->halo_err.p, sizeof(unsigned), hipMemcpyDeviceToHost, ctx->stream),
            "read halo flags");
    SPG_HIP(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return SPG_OK;
}

}  // namespace

spg_status device_halo_operands(spg_context *ctx, const spg_csr_view *A, const spg_csr_view *B, int32_t b_first_row,
                                const spg_halo_view *halo, int32_t ndir, Operands &op, int64_t &b_span) {
    hipStream_t st = ctx->stream;
    const int32_t hn = halo ? halo->n_rows : 0;
    if (hn < 0 || (hn > 0 && (!halo->global_rows || !halo->row_ptr)))
        return set_err(ctx, SPG_EINVAL, "malformed halo view");
    int64_t hnnz = 0;
    if (hn > 0) {  // on the context's stream: ordered after work that filled the halo there
        SPG_HIP(hipMemcpyAsync(&hnnz, halo->row_ptr + hn, sizeof(int64_t), hipMemcpyDeviceToHost, st),
                "read halo size");
        SPG_HIP(hipStreamSynchronize(st), "hipStreamSynchronize");
    }
    if (hnnz < 0) return set_err(ctx, SPG_EINVAL, "malformed halo view");
    if (hnnz > 0 && (!halo->col_idx || !halo->values)) return set_err(ctx, SPG_EINVAL, "malformed halo view");
    b_span = B->nnz + hnnz;
    const size_t tot = (size_t)b_span;
    SPG_HIP(ctx->b_col.ensure(tot * 4, st), "hipMalloc(B col)");
    SPG_HIP(ctx->b_val.ensure(tot * 8, st), "hipMalloc(B val)");
    SPG_HIP(ctx->dir_start.ensure((size_t)ndir * 8, st), "hipMalloc(dir)");
    SPG_HIP(ctx->dir_end.ensure((size_t)ndir * 8, st), "hipMalloc(dir)");
    SPG_HIP(ctx->halo_err.ensure(sizeof(unsigned), st), "hipMalloc(halo flags)");
    // Entries [B_local | halo], device to device.
    if (B->nnz) {
        SPG_HIP(hipMemcpyAsync(ctx->b_col.p, B->col_idx, (size_t)B->nnz * 4, hipMemcpyDeviceToDevice, st), "copy B col");
        SPG_HIP(hipMemcpyAsync(ctx->b_val.p, B->values, (size_t)B->nnz * 8, hipMemcpyDeviceToDevice, st), "copy B val");
    }
    if (hnnz) {
        SPG_HIP(hipMemcpyAsync(ctx->b_col.as<int32_t>() + B->nnz, halo->col_idx, (size_t)hnnz * 4,
                               hipMemcpyDeviceToDevice, st), "copy halo col");
        SPG_HIP(hipMemcpyAsync(ctx->b_val.as<double>() + B->nnz, halo->values, (size_t)hnnz * 8,
                               hipMemcpyDeviceToDevice, st), "copy halo val");
    }
    int64_t *ds = ctx->dir_start.as<int64_t>(), *de = ctx->dir_end.as<int64_t>();
    unsigned *err = ctx->halo_err.as<unsigned>();
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, st, err);
    if (ndir > 0) {
        hipLaunchKernelGGL(k_fill_i64, dim3(grid_for(ndir, 1024)), dim3(256), 0, st, ds, (int64_t)ndir, (int64_t)-1);
        hipLaunchKernelGGL(k_fill_i64, dim3(grid_for(ndir, 1024)), dim3(256), 0, st, de, (int64_t)ndir, (int64_t)-1);
    }
    if (B->rows > 0)
        hipLaunchKernelGGL(k_dir_local, dim3(grid_for(B->rows, 256)), dim3(256), 0, st, B->row_ptr, B->rows, b_first_row,
                           ds, de);
    if (hn > 0)
        hipLaunchKernelGGL(k_dir_halo, dim3(grid_for(hn, 256)), dim3(256), 0, st, halo->global_rows, halo->row_ptr, hn,
                           b_first_row, B->rows, ndir, B->nnz, ds, de, err);
    SPG_HIP(hipGetLastError(), "kernel launch");
    unsigned flags = 0;
    spg_status s = read_flags(ctx, flags);
    if (s != SPG_OK) return s;
    if (flags & kHaloRange) return set_err(ctx, SPG_EINVAL, "halo row outside A's column range");
    if (flags & kHaloDuplicate) return set_err(ctx, SPG_EINVAL, "duplicate halo row");
    if (flags & kHaloBadPtr) return set_err(ctx, SPG_EINVAL, "halo row_ptr is negative or not nondecreasing");
    op.ndir = ndir;
    op.bstart = ds;
    op.bend = de;
    op.bcol = ctx->b_col.as<int32_t>();
    op.bval = ctx->b_val.as<double>();
    return SPG_OK;
}

}  // namespace spg

using namespace spg;

// Library-owned halo rows in HBM: global row of each slot, slot offsets,
// entries.
struct spg_device_halo {
    int device = 0;
    int32_t n = 0;
    int64_t nnz = 0;
    void *rows = nullptr, *ptr = nullptr, *col = nullptr, *val = nullptr;
};

namespace {

spg_status halo_alloc(spg_context *ctx, spg_device_halo *h, int32_t n, int64_t nnz) {
    h->device = ctx->device;
    h->n = n;
    h->nnz = nnz;
    hipError_t e = hipMalloc(&h->rows, n ? (size_t)n * 4 : 4);
    if (e == hipSuccess) e = hipMalloc(&h->ptr, ((size_t)n + 1) * 8);
    if (e == hipSuccess) e = hipMalloc(&h->col, nnz ? (size_t)nnz * 4 : 4);
    if (e == hipSuccess) e = hipMalloc(&h->val, nnz ? (size_t)nnz * 8 : 8);
    return e == hipSuccess ? SPG_OK : hip_err(ctx, e, "allocate halo rows");
}

}  // namespace

extern "C" {

spg_status spg_halo_upload(spg_context *ctx, const spg_halo_view *host, spg_device_halo **out) {
    if (!ctx) return SPG_EINVAL;
    if (!host || !out) return set_err(ctx, SPG_EINVAL, "null argument");
    *out = nullptr;
    if (host->location != SPG_HOST) return set_err(ctx, SPG_EINVAL, "upload takes a host halo view");
    if (host->n_rows < 0 || (host->n_rows > 0 && (!host->global_rows || !host->row_ptr)))
        return set_err(ctx, SPG_EINVAL, "malformed halo view");
    const int32_t n = host->n_rows;
    const int64_t nnz = n ? host->row_ptr[n] : 0;
    if (nnz < 0 || (nnz > 0 && (!host->col_idx || !host->values))) return set_err(ctx, SPG_EINVAL, "malformed halo view");
    SPG_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    spg_device_halo *h = new (std::nothrow) spg_device_halo();
    if (!h) return set_err(ctx, SPG_ENOMEM, "out of memory");
    spg_status s = halo_alloc(ctx, h, n, nnz);
    hipError_t e = hipSuccess;
    if (s == SPG_OK && n) e = hipMemcpy(h->rows, host->global_rows, (size_t)n * 4, hipMemcpyHostToDevice);
    if (s == SPG_OK && e == hipSuccess) {
        if (n) e = hipMemcpy(h->ptr, host->row_ptr, ((size_t)n + 1) * 8, hipMemcpyHostToDevice);
        else e = hipMemset(h->ptr, 0, 8);
    }
    if (s == SPG_OK && e == hipSuccess && nnz) e = hipMemcpy(h->col, host->col_idx, (size_t)nnz * 4, hipMemcpyHostToDevice);
    if (s == SPG_OK && e == hipSuccess && nnz) e = hipMemcpy(h->val, host->values, (size_t)nnz * 8, hipMemcpyHostToDevice);
    if (s == SPG_OK && e != hipSuccess) s = hip_err(ctx, e, "upload halo rows");
    if (s != SPG_OK) {
        spg_device_halo_free(h);
        return s;
    }
    *out = h;
    return SPG_OK;
}

spg_status spg_halo_pack(spg_context *ctx, const spg_csr_view *B_local, int32_t first_row, const int32_t *requests,
                         int32_t n, int32_t requests_location, spg_device_halo **out) {
    if (!ctx) return SPG_EINVAL;
    if (!B_local || !out || n < 0 || (n > 0 && !requests)) return set_err(ctx, SPG_EINVAL, "null argument");
    *out = nullptr;
    if (B_local->location != SPG_DEVICE) return set_err(ctx, SPG_EINVAL, "packing takes a device view of B_local");
    spg_status s = check_host_csr(ctx, B_local, "B_local");
    if (s != SPG_OK) return s;
    SPG_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    hipStream_t st = ctx->stream;
    spg_device_halo *h = new (std::nothrow) spg_device_halo();
    if (!h) return set_err(ctx, SPG_ENOMEM, "out of memory");
    // Rows, offsets, then entries once the total is known.
    h->device = ctx->device;
    h->n = n;
    hipError_t e = hipMalloc(&h->rows, n ? (size_t)n * 4 : 4);
    if (e == hipSuccess) e = hipMalloc(&h->ptr, ((size_t)n + 1) * 8);
    if (e != hipSuccess) {
        spg_device_halo_free(h);
        return hip_err(ctx, e, "allocate halo rows");
    }
    if (n) {
        e = hipMemcpyAsync(h->rows, requests, (size_t)n * 4,
                           requests_location == SPG_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st);
        if (e != hipSuccess) {
            spg_device_halo_free(h);
            return hip_err(ctx, e, "copy requests");
        }
    }
    if ((s = [&]() -> spg_status {
             SPG_HIP(ctx->halo_cnt.ensure(((size_t)n + 1) * 8, st), "hipMalloc(counts)");
             SPG_HIP(ctx->halo_tmp.ensure(scan_tmp_elems(n) * 8, st), "hipMalloc(scan)");
             SPG_HIP(ctx->halo_err.ensure(sizeof(unsigned), st), "hipMalloc(halo flags)");
             return SPG_OK;
         }()) != SPG_OK) {
        spg_device_halo_free(h);
        return s;
    }
    unsigned *err = ctx->halo_err.as<unsigned>();
    int64_t *cnt = ctx->halo_cnt.as<int64_t>();
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, st, err);
    if (n)
        hipLaunchKernelGGL(k_pack_counts, dim3(grid_for(n, 256)), dim3(256), 0, st, B_local->row_ptr, B_local->rows,
                           first_row, static_cast<const int32_t *>(h->rows), n, cnt, err);
    launch_scan(cnt, static_cast<int64_t *>(h->ptr), n, ctx->halo_tmp.as<int64_t>(), nullptr, st);
    unsigned flags = 0;
    int64_t total = 0;
    if ((e = hipGetLastError()) == hipSuccess) e = hipMemcpyAsync(&total, static_cast<int64_t *>(h->ptr) + n, 8,
                                                                   hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) s = read_flags(ctx, flags);
    else s = hip_err(ctx, e, "pack halo counts");
    if (s == SPG_OK && (flags & kPackNotLocal)) s = set_err(ctx, SPG_EMISSING_ROW, "received a request for a non-local sparse row of B");
    if (s == SPG_OK) {
        h->nnz = total;
        e = hipMalloc(&h->col, total ? (size_t)total * 4 : 4);
        if (e == hipSuccess) e = hipMalloc(&h->val, total ? (size_t)total * 8 : 8);
        if (e == hipSuccess && n)
            hipLaunchKernelGGL(k_pack_rows, dim3(grid_for((int64_t)n * 64, 256)), dim3(256), 0, st, B_local->row_ptr,
                               B_local->col_idx, B_local->values, first_row, static_cast<const int32_t *>(h->rows), n,
                               static_cast<const int64_t *>(h->ptr), static_cast<int32_t *>(h->col),
                               static_cast<double *>(h->val));
        if (e == hipSuccess) e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) s = hip_err(ctx, e, "pack halo rows");
    }
    if (s != SPG_OK) {
        spg_device_halo_free(h);
        return s;
    }
    *out = h;
    return SPG_OK;
}

spg_status spg_device_halo_view(const spg_device_halo *h, spg_halo_view *out) {
    if (!h || !out) return SPG_EINVAL;
    out->n_rows = h->n;
    out->global_rows = static_cast<const int32_t *>(h->rows);
    out->row_ptr = static_cast<const int64_t *>(h->ptr);
    out->col_idx = static_cast<const int32_t *>(h->col);
    out->values = static_cast<const double *>(h->val);
    out->location = SPG_DEVICE;
    return SPG_OK;
}

spg_status spg_device_halo_free(spg_device_halo *h) {
    if (!h) return SPG_OK;
    (void)hipSetDevice(h->device);
    if (h->rows) hipFree(h->rows);
    if (h->ptr) hipFree(h->ptr);
    if (h->col) hipFree(h->col);
    if (h->val) hipFree(h->val);
    delete h;
    return SPG_OK;
}

}  // extern "C"
