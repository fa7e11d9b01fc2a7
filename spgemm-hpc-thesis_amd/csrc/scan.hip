// scan.hip -- int64 exclusive scan (C row offsets) and row compaction.
#include "device_common.hpp"

namespace spg {
namespace {

using namespace dev;

constexpr int kScanTpb = 256;
constexpr int kScanIpt = 8;
constexpr int kScanTile = kScanTpb * kScanIpt;

__global__ __launch_bounds__(kScanTpb) void k_scan_reduce(const int64_t *in, int64_t n, int64_t *partial) {
    __shared__ int64_t s_ws[kScanTpb / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    int64_t sum = 0;
#pragma unroll
    for (int j = 0; j < kScanIpt; ++j) {
        const int64_t i = base + (int64_t)j * kScanTpb + threadIdx.x;
        if (i < n) sum += in[i];
    }
    const int64_t total = block_sum<kScanTpb>(sum, s_ws);
    if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanTpb) void k_scan_partials(int64_t *partial, int64_t nparts, int64_t *grand,
                                                            unsigned long long *total_dst) {
    __shared__ int64_t s_ws[kScanTpb / 64];
    int64_t carry = 0;
    for (int64_t b0 = 0; b0 < nparts; b0 += kScanTpb) {
        const int64_t i = b0 + threadIdx.x;
        const int64_t v = i < nparts ? partial[i] : 0;
        int64_t tot;
        const int64_t ex = block_excl_scan<kScanTpb>(v, s_ws, tot);
        if (i < nparts) partial[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) {
        *grand = carry;
        if (total_dst) *total_dst = (unsigned long long)carry;
    }
}

__global__ __launch_bounds__(kScanTpb) void k_scan_down(const int64_t *in, int64_t n, const int64_t *partial,
                                                        int64_t *out, const int64_t *grand) {
    __shared__ int64_t s_ws[kScanTpb / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanIpt;
    int64_t v[kScanIpt];
    int64_t sum = 0;
#pragma unroll
    for (int j = 0; j < kScanIpt; ++j) {
        const int64_t i = base + j;
        v[j] = i < n ? in[i] : 0;
        sum += v[j];
    }
    int64_t tot;
    int64_t run = partial[blockIdx.x] + block_excl_scan<kScanTpb>(sum, s_ws, tot);
#pragma unroll
    for (int j = 0; j < kScanIpt; ++j) {
        const int64_t i = base + j;
        if (i < n) out[i] = run;
        run += v[j];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = *grand;
}

// Adds the sum of in[0, n) to *dst: eight independent loads in flight per
// thread, one atomic per workgroup.
constexpr int kSumUnroll = 8;
__global__ __launch_bounds__(kScanTpb) void k_sum(const int64_t *in, int64_t n, unsigned long long *dst) {
    __shared__ int64_t s_ws[kScanTpb / 64];
    const int64_t stride = (int64_t)gridDim.x * kScanTpb;
    int64_t acc[kSumUnroll] = {};
    for (int64_t i0 = (int64_t)blockIdx.x * kScanTpb + threadIdx.x; i0 < n; i0 += stride * kSumUnroll) {
#pragma unroll
        for (int u = 0; u < kSumUnroll; ++u) {
            const int64_t i = i0 + u * stride;
            acc[u] += i < n ? in[i] : 0;
        }
    }
    int64_t v = 0;
#pragma unroll
    for (int u = 0; u < kSumUnroll; ++u) v += acc[u];
    v = block_sum<kScanTpb>(v, s_ws);
    if (threadIdx.x == 0 && v != 0) atomicAdd(dst, (unsigned long long)v);
}

// One wave per row (only when exact zeros were dropped).
__global__ __launch_bounds__(256) void k_compact(int32_t rows, const int64_t *rp_in, const int32_t *col_in,
                                                 const double *val_in, const int64_t *rp_out, int32_t *col_out,
                                                 double *val_out) {
    const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int nw = (int)((gridDim.x * blockDim.x) >> 6);
    for (int r = wave; r < rows; r += nw) {
        const int64_t src = rp_in[r];
        const int64_t dst = rp_out[r];
        const int64_t n = rp_out[r + 1] - dst;
        for (int64_t q = lane_id(); q < n; q += 64) {
            col_out[dst + q] = col_in[src + q];
            val_out[dst + q] = val_in[src + q];
        }
    }
}

}  // namespace

size_t scan_tmp_elems(int64_t n) {
    const int64_t parts = (n + kScanTile - 1) / kScanTile;
    return (size_t)(parts + 2);
}

void launch_scan(const int64_t *in, int64_t *out, int64_t n, int64_t *tmp, unsigned long long *total_dst,
                 hipStream_t stream) {
    const int64_t parts = (n + kScanTile - 1) / kScanTile;
    int64_t *grand = tmp + parts;
    if (parts > 0) hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)parts), dim3(kScanTpb), 0, stream, in, n, tmp);
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(kScanTpb), 0, stream, tmp, parts, grand, total_dst);
    const int64_t launch_parts = parts > 0 ? parts : 1;
    hipLaunchKernelGGL(k_scan_down, dim3((unsigned)launch_parts), dim3(kScanTpb), 0, stream, in, n, tmp, out, grand);
}

void launch_sum(const int64_t *in, int64_t n, unsigned long long *dst, hipStream_t stream) {
    if (n <= 0) return;
    int64_t blocks = (n + (int64_t)kScanTpb * kSumUnroll - 1) / ((int64_t)kScanTpb * kSumUnroll);
    if (blocks > 256) blocks = 256;
    hipLaunchKernelGGL(k_sum, dim3((unsigned)blocks), dim3(kScanTpb), 0, stream, in, n, dst);
}

void launch_compact(int32_t rows, const int64_t *rowptr_in, const int32_t *col_in, const double *val_in,
                    const int64_t *rowptr_out, int32_t *col_out, double *val_out, hipStream_t stream) {
    if (rows <= 0) return;
    int blocks = (int)(((int64_t)rows * 64 + 255) / 256);
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_compact, dim3(blocks), dim3(256), 0, stream, rows, rowptr_in, col_in, val_in, rowptr_out,
                       col_out, val_out);
}

}  // namespace spg
