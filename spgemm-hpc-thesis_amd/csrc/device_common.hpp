// device_common.hpp -- wave/block primitives shared by the gfx950 kernels.
#pragma once

#include "internal.hpp"

#include <climits>

namespace spg {
namespace dev {

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

__device__ __forceinline__ unsigned long long lanemask_lt() { return (1ull << lane_id()) - 1ull; }

// Orders this wave's LDS accesses around a phase change (waves of a block
// that work independently on their own LDS slices need no s_barrier).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        T t = __shfl_up(v, d, 64);
        if (lane_id() >= d) v += t;
    }
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v = min(v, __shfl_xor(v, d, 64));
    return v;
}

__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v = max(v, __shfl_xor(v, d, 64));
    return v;
}

// Exclusive scan across a TPB-thread block; `ws` holds TPB/64 elements.
// Has barriers when TPB > 64 and must then be reached by every thread.
template <int TPB, typename T>
__device__ __forceinline__ T block_excl_scan(T v, T *ws, T &total) {
    constexpr int NW = TPB / 64;
    const T inc = wave_incl_scan(v);
    if constexpr (NW == 1) {
        total = __shfl(inc, 63, 64);
        return inc - v;
    } else {
        const int w = threadIdx.x >> 6;
        if (lane_id() == 63) ws[w] = inc;
        __syncthreads();
        T before = 0, all = 0;
#pragma unroll
        for (int i = 0; i < NW; ++i) {
            const T x = ws[i];
            if (i < w) before += x;
            all += x;
        }
        __syncthreads();
        total = all;
        return before + inc - v;
    }
}

template <int TPB, typename T>
__device__ __forceinline__ T block_sum(T v, T *ws) {
    T total;
    block_excl_scan<TPB>(v, ws, total);
    return total;
}

template <int TPB>
__device__ __forceinline__ int block_max(int v, int *ws) {
    constexpr int NW = TPB / 64;
    v = wave_max(v);
    if constexpr (NW > 1) {
        const int w = threadIdx.x >> 6;
        if (lane_id() == 0) ws[w] = v;
        __syncthreads();
        int m = INT_MIN;
#pragma unroll
        for (int i = 0; i < NW; ++i) m = max(m, ws[i]);
        __syncthreads();
        v = m;
    }
    return v;
}

template <int TPB>
__device__ __forceinline__ void block_minmax(int &mn, int &mx, int *ws) {
    constexpr int NW = TPB / 64;
    mn = wave_min(mn);
    mx = wave_max(mx);
    if constexpr (NW > 1) {
        const int w = threadIdx.x >> 6;
        if (lane_id() == 0) {
            ws[w] = mn;
            ws[NW + w] = mx;
        }
        __syncthreads();
        int a = INT_MAX, b = INT_MIN;
#pragma unroll
        for (int i = 0; i < NW; ++i) {
            a = min(a, ws[i]);
            b = max(b, ws[NW + i]);
        }
        __syncthreads();
        mn = a;
        mx = b;
    }
}

// B row k through the directory; flags record a bad column or a missing row.
__device__ __forceinline__ void brow(const Operands &op, int32_t k, int64_t &s, int64_t &e, unsigned &flags) {
    if (k < 0 || k >= op.ndir) {
        flags |= kErrRange;
        s = e = 0;
        return;
    }
    s = op.bstart[k];
    e = op.bend[k];
    if (s < 0) {
        flags |= kErrMissing;
        s = e = 0;
    } else if (e < s) {
        e = s;
    }
}

// Largest e in [0, n) with off[e] <= s, given off[0] <= s (off nondecreasing).
template <typename T>
__device__ __forceinline__ int upper_entry(const T *off, int n, T s) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= s) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ int classify(int64_t ub, int64_t na) {
    if (ub == 0) return kBinEmpty;
    if (ub <= kCapW1 && na <= kCapW1) return kBinW1;
    if (ub <= kCapW2 && na <= kCapW2) return kBinW2;
    if (ub <= kCapW3 && na <= kCapW3) return kBinW3;
    if (ub <= kCapBlock && na <= kCapBlock) return kBinBlock;
    return kBinLarge;
}

// Presence bitmaps of the large-row path are updated with L2 atomics by many
// threads; plain loads could hit a stale L1 line, so reads and clears go to L2.
__device__ __forceinline__ uint32_t ld_bits(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_bits(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Rows of one bin: list and count (filled by the analyze pass).
__device__ __forceinline__ const int32_t *bin_list(const Workspace &ws, int bin) {
    return ws.bin_rows + (size_t)bin * (size_t)ws.list_stride;
}

template <typename K>
int occupancy_grid(K kernel, int tpb, int cus, long long want) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, tpb, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    long long g = (long long)per_cu * cus;
    if (want > 0 && want < g) g = want;
    return (int)(g < 1 ? 1 : g);
}

}  // namespace dev
}  // namespace spg
