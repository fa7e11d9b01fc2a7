// symbolic.hip -- exact structural count (distinct output columns) per row.
//
// The count does not depend on summation order, so small and medium rows use
// an open-addressing hash set of column indices in LDS (one compare-and-swap
// per product, table at least twice the row's product bound, so probing always
// ends).  A entries are expanded "entry-major": with L the power of two >= the
// longest referenced B row, each group of L lanes streams one B row, so there
// is no per-product search and reads of a B row are contiguous.
// Large rows set bits in a per-workgroup column bitmap in HBM (window by window)
// and count the bits they newly set.
#include "device_common.hpp"

namespace spg {
namespace {

using namespace dev;

constexpr int32_t kEmptyKey = INT_MAX;  // never a valid column (col < cols <= INT_MAX)

__device__ __forceinline__ uint32_t hash_col(int32_t c) { return (uint32_t)c * 2654435761u; }

template <int TPB, int CAP>
__global__ __launch_bounds__(TPB) void k_symbolic_hash(Operands op, Workspace ws, int binid) {
    constexpr int TMAX = 2 * CAP;
    __shared__ int32_t s_tab[TMAX];
    __shared__ int64_t s_bst[CAP];
    __shared__ int32_t s_len[CAP];
    __shared__ int s_wsi[TPB / 64 + 1];
    const int tid = threadIdx.x;
    for (int i = tid; i < TMAX; i += TPB) s_tab[i] = kEmptyKey;
    __syncthreads();
    const int32_t *list = bin_list(ws, binid);
    const int count = ws.ctrl->bin_count[binid];
    unsigned flags = 0;
    for (int i = (int)blockIdx.x; i < count; i += (int)gridDim.x) {
        const int row = list[i];
        const int64_t a0 = op.aptr[row];
        const int na = (int)(op.aptr[row + 1] - a0);  // <= CAP by binning
        int mylen = 0, mymax = 0;
        for (int e = tid; e < na; e += TPB) {
            const int64_t s = __builtin_nontemporal_load(&ws.ebst[a0 + e]);
            const int len = __builtin_nontemporal_load(&ws.elen[a0 + e]);
            s_bst[e] = s;
            s_len[e] = len;
            mylen += len;
            mymax = max(mymax, len);
        }
        const int P = block_sum<TPB>(mylen, s_wsi);
        const int L = block_max<TPB>(mymax, s_wsi);
        if (P > CAP) {  // cannot happen: binning used the same lengths (uniform across the block)
            flags |= kErrInternal;
            __syncthreads();
            continue;
        }
        int T = 64;
        while (T < 2 * P && T < TMAX) T <<= 1;
        const int logT = 31 - __clz(T);
        __syncthreads();
        int fresh = 0;
        // T >= 2P > distinct columns, so an empty slot is always found; the
        // probe bound only makes termination unconditional.
        auto insert = [&](int32_t c) {
            uint32_t h = hash_col(c) >> (32 - logT);
            for (int probe = 0; probe < T; ++probe) {
                const int32_t old = atomicCAS(&s_tab[h], kEmptyKey, c);
                if (old == kEmptyKey) {
                    ++fresh;
                    return;
                }
                if (old == c) return;
                h = (h + 1u) & (uint32_t)(T - 1);
            }
            flags |= kErrInternal;
        };
        if (L <= TPB) {
            int Lp = 1;
            while (Lp < L) Lp <<= 1;
            const int logL = 31 - __clz(Lp);
            const int G = TPB >> logL;
            const int sub = tid & (Lp - 1);
            for (int e0 = 0; e0 < na; e0 += G) {
                const int e = e0 + (tid >> logL);
                if (e < na && sub < s_len[e]) insert(op.bcol[s_bst[e] + sub]);
            }
        } else {
            for (int e = 0; e < na; ++e) {
                const int len = s_len[e];
                const int64_t s = s_bst[e];
                for (int t = tid; t < len; t += TPB) insert(op.bcol[s + t]);
            }
        }
        const int total = block_sum<TPB>(fresh, s_wsi);
        if (tid == 0) ws.snnz[row] = total;
        __syncthreads();
        for (int q = tid; q < T; q += TPB) s_tab[q] = kEmptyKey;
        __syncthreads();
    }
    if (flags) atomicOr(&ws.ctrl->flags, (unsigned long long)flags);
}

template <int TPB, int CAP>
__device__ __forceinline__ int64_t large_batch_meta(const Operands &op, const Workspace &ws, int64_t e0, int nbt,
                                                    int64_t *s_bst, double *s_av, int64_t *s_off, int64_t *s_ws64,
                                                    bool want_values) {
    const int tid = threadIdx.x;
    const int ipt = (nbt + TPB - 1) / TPB;
    const int f = tid * ipt;
    int64_t mylen = 0;
    for (int j = 0; j < ipt; ++j) {
        const int e = f + j;
        if (e < nbt) {
            const int64_t len = ws.elen[e0 + e];
            s_bst[e] = ws.ebst[e0 + e];
            if (want_values) s_av[e] = op.aval[e0 + e];
            s_off[e] = len;
            mylen += len;
        }
    }
    int64_t Pb;
    int64_t base = block_excl_scan<TPB>(mylen, s_ws64, Pb);
    for (int j = 0; j < ipt; ++j) {
        const int e = f + j;
        if (e < nbt) {
            const int64_t l = s_off[e];
            s_off[e] = base;
            base += l;
        }
    }
    __syncthreads();
    return Pb;
}

template <int TPB, int CAP>
__global__ __launch_bounds__(TPB) void k_large_symbolic(Operands op, Workspace ws) {
    __shared__ int64_t s_bst[CAP];
    __shared__ int64_t s_off[CAP + 1];
    __shared__ int64_t s_ws64[TPB / 64 + 1];
    const int tid = threadIdx.x;
    const int32_t *list = bin_list(ws, kBinLarge);
    const int count = ws.ctrl->bin_count[kBinLarge];
    const int64_t W = ws.lwidth;
    uint32_t *bits = ws.lbits + (size_t)blockIdx.x * (size_t)(W / 32);
    unsigned flags = 0;
    for (int i = (int)blockIdx.x; i < count; i += (int)gridDim.x) {
        const int row = list[i];
        const int64_t a0 = op.aptr[row], a1 = op.aptr[row + 1];
        int64_t cnt = 0;
        for (int64_t w0 = 0; w0 < op.bcols; w0 += W) {
            const int64_t wend = min(w0 + W, (int64_t)op.bcols);
            for (int64_t e0 = a0; e0 < a1; e0 += CAP) {
                const int nbt = (int)min((int64_t)CAP, a1 - e0);
                const int64_t Pb = large_batch_meta<TPB, CAP>(op, ws, e0, nbt, s_bst, nullptr, s_off, s_ws64, false);
                for (int64_t s = tid; s < Pb; s += TPB) {
                    const int e = upper_entry(s_off, nbt, s);
                    const int64_t c = op.bcol[s_bst[e] + (s - s_off[e])];
                    if (c >= w0 && c < wend) {
                        const uint32_t rel = (uint32_t)(c - w0);
                        const uint32_t bit = 1u << (rel & 31u);
                        const uint32_t old = atomicOr(&bits[rel >> 5], bit);
                        if (!(old & bit)) ++cnt;
                    }
                }
                __syncthreads();  // batch metadata is overwritten next
            }
            const int64_t nwords = (wend - w0 + 31) / 32;
            for (int64_t q = tid; q < nwords; q += TPB) st_bits(&bits[q], 0u);
            __threadfence();
            __syncthreads();
        }
        const int64_t total = block_sum<TPB>(cnt, s_ws64);
        if (tid == 0) ws.snnz[row] = total;
    }
    if (flags) atomicOr(&ws.ctrl->flags, (unsigned long long)flags);
}

template <int TPB, int CAP>
void launch_hash(const Operands &op, const Workspace &ws, int bin, int cus, hipStream_t stream) {
    auto k = k_symbolic_hash<TPB, CAP>;
    const int g = occupancy_grid(k, TPB, cus, op.arows);
    hipLaunchKernelGGL(k, dim3(g), dim3(TPB), 0, stream, op, ws, bin);
}

}  // namespace

void launch_symbolic(const Operands &op, const Workspace &ws, int cus, hipStream_t stream) {
    if (op.arows <= 0) return;
    launch_hash<64, kCapW1>(op, ws, kBinW1, cus, stream);
    launch_hash<64, kCapW2>(op, ws, kBinW2, cus, stream);
    launch_hash<64, kCapW3>(op, ws, kBinW3, cus, stream);
    launch_hash<kTpbBlock, kCapBlock>(op, ws, kBinBlock, cus, stream);
    if (ws.lblocks > 0)
        hipLaunchKernelGGL((k_large_symbolic<kTpbLarge, kCapLarge>), dim3(ws.lblocks), dim3(kTpbLarge), 0, stream, op,
                           ws);
}

}  // namespace spg
