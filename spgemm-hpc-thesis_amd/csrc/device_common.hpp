// device_common.hpp -- wave/block primitives shared by the gfx950 kernels.
#pragma once

#include "internal.hpp"

#include <climits>
#include <type_traits>

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

// DPP forms of the wave scans and reductions (no LDS traffic): row_shr 1, 2,
// 4, 8 inside each 16-lane row, then row_bcast 15 and 31 carry the row totals
// upward (row masks 0xa and 0xc).  Lanes whose source is outside the row read
// the identity.  Every lane of the wave must be active.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ int dpp_i(int identity, int v) {
    return __builtin_amdgcn_update_dpp(identity, v, CTRL, ROWMASK, 0xf, false);
}
// The same move of a 64-bit value (both halves from the same lane).
template <int CTRL, int ROWMASK>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t identity, uint64_t v) {
    const uint32_t lo = (uint32_t)dpp_i<CTRL, ROWMASK>((int)(uint32_t)identity, (int)(uint32_t)v);
    const uint32_t hi = (uint32_t)dpp_i<CTRL, ROWMASK>((int)(uint32_t)(identity >> 32), (int)(uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ int wave_incl_sum_dpp(int v) {
    v += dpp_i<0x111, 0xf>(0, v);
    v += dpp_i<0x112, 0xf>(0, v);
    v += dpp_i<0x114, 0xf>(0, v);
    v += dpp_i<0x118, 0xf>(0, v);
    v += dpp_i<0x142, 0xa>(0, v);
    v += dpp_i<0x143, 0xc>(0, v);
    return v;
}
__device__ __forceinline__ uint64_t wave_incl_sum_dpp64(uint64_t v) {
    v += dpp_u64<0x111, 0xf>(0, v);
    v += dpp_u64<0x112, 0xf>(0, v);
    v += dpp_u64<0x114, 0xf>(0, v);
    v += dpp_u64<0x118, 0xf>(0, v);
    v += dpp_u64<0x142, 0xa>(0, v);
    v += dpp_u64<0x143, 0xc>(0, v);
    return v;
}

// Lane 63 of an inclusive scan holds the wave's result (a scalar register).
__device__ __forceinline__ int wave_last(int v) { return __builtin_amdgcn_readlane(v, 63); }
__device__ __forceinline__ uint64_t wave_last64(uint64_t v) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, 63);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), 63);
    return ((uint64_t)hi << 32) | lo;
}

template <typename T>
constexpr bool kDppInt = std::is_integral<T>::value && (sizeof(T) == 4 || sizeof(T) == 8);

// Inclusive scan over the wave's lanes (all active).  Integers take DPP moves;
// other types (none in the hot paths) take lane shuffles.
template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v) {
    if constexpr (kDppInt<T> && sizeof(T) == 4) {
        return (T)wave_incl_sum_dpp((int)v);
    } else if constexpr (kDppInt<T>) {
        return (T)wave_incl_sum_dpp64((uint64_t)v);
    } else {
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            T t = __shfl_up(v, d, 64);
            if (lane_id() >= d) v += t;
        }
        return v;
    }
}

// Lane 63's value, in every lane.
template <typename T>
__device__ __forceinline__ T wave_bcast_last(T v) {
    if constexpr (kDppInt<T> && sizeof(T) == 4) return (T)wave_last((int)v);
    else if constexpr (kDppInt<T>) return (T)wave_last64((uint64_t)v);
    else return __shfl(v, 63, 64);
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
    if constexpr (kDppInt<T>) {
        return wave_bcast_last(wave_incl_scan(v));
    } else {
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
        return v;
    }
}

__device__ __forceinline__ int wave_max_dpp(int v) {
    v = max(v, dpp_i<0x111, 0xf>(INT_MIN, v));
    v = max(v, dpp_i<0x112, 0xf>(INT_MIN, v));
    v = max(v, dpp_i<0x114, 0xf>(INT_MIN, v));
    v = max(v, dpp_i<0x118, 0xf>(INT_MIN, v));
    v = max(v, dpp_i<0x142, 0xa>(INT_MIN, v));
    v = max(v, dpp_i<0x143, 0xc>(INT_MIN, v));
    return wave_last(v);
}

// Inclusive prefix maximum over the lanes (lane l gets max of lanes 0..l).
__device__ __forceinline__ int wave_incl_max_dpp(int v) {
    v = max(v, dpp_i<0x111, 0xf>(INT_MIN, v));
    v = max(v, dpp_i<0x112, 0xf>(INT_MIN, v));
    v = max(v, dpp_i<0x114, 0xf>(INT_MIN, v));
    v = max(v, dpp_i<0x118, 0xf>(INT_MIN, v));
    v = max(v, dpp_i<0x142, 0xa>(INT_MIN, v));
    v = max(v, dpp_i<0x143, 0xc>(INT_MIN, v));
    return v;
}

__device__ __forceinline__ int wave_min_dpp(int v) {
    v = min(v, dpp_i<0x111, 0xf>(INT_MAX, v));
    v = min(v, dpp_i<0x112, 0xf>(INT_MAX, v));
    v = min(v, dpp_i<0x114, 0xf>(INT_MAX, v));
    v = min(v, dpp_i<0x118, 0xf>(INT_MAX, v));
    v = min(v, dpp_i<0x142, 0xa>(INT_MAX, v));
    v = min(v, dpp_i<0x143, 0xc>(INT_MAX, v));
    return wave_last(v);
}

__device__ __forceinline__ int wave_min(int v) { return wave_min_dpp(v); }

// OR over the wave's lanes (all active), in every lane.
__device__ __forceinline__ unsigned wave_or(unsigned v) {
    v |= (unsigned)dpp_i<0x111, 0xf>(0, (int)v);
    v |= (unsigned)dpp_i<0x112, 0xf>(0, (int)v);
    v |= (unsigned)dpp_i<0x114, 0xf>(0, (int)v);
    v |= (unsigned)dpp_i<0x118, 0xf>(0, (int)v);
    v |= (unsigned)dpp_i<0x142, 0xa>(0, (int)v);
    v |= (unsigned)dpp_i<0x143, 0xc>(0, (int)v);
    return (unsigned)wave_last((int)v);
}

// Lane k's 64-bit value (k uniform), in every lane.
__device__ __forceinline__ int64_t lane_value64(int64_t v, int k) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, k);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)v >> 32), k);
    return (int64_t)(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ int wave_max(int v) { return wave_max_dpp(v); }

// Exclusive scan across a TPB-thread block; `ws` holds TPB/64 elements.
// Has barriers when TPB > 64 and must then be reached by every thread.
template <int TPB, typename T>
__device__ __forceinline__ T block_excl_scan(T v, T *ws, T &total) {
    constexpr int NW = TPB / 64;
    const T inc = wave_incl_scan(v);
    if constexpr (NW == 1) {
        total = wave_bcast_last(inc);
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

// Exclusive scan of s[0, n) in place, n a multiple of TPB: each thread sums
// n / TPB consecutive counters, one block scan orders the threads.  Returns
// the total; the written offsets are visible after the caller's next barrier.
template <int TPB, typename T>
__device__ __forceinline__ T block_scan_runs(T *s, int n, T *ws) {
    const int per = n / TPB;
    T *mine = s + threadIdx.x * per;
    T loc = 0;
    for (int k = 0; k < per; ++k) loc += mine[k];
    T total;
    T run = block_excl_scan<TPB>(loc, ws, total);
    for (int k = 0; k < per; ++k) {
        const T c = mine[k];
        mine[k] = run;
        run += c;
    }
    return total;
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

// Block-uniform value held in a scalar register (the value must already be
// equal in every lane; this lets the compiler branch on it with scalar code).
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int64_t uni64(int64_t v) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
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

// upper_entry for ITEMS queries at once (off[0] <= s[j] for valid queries;
// others get some e in [0, n)).  The searches advance in lockstep for the
// ceil(log2 n) steps any of them needs, so a thread's LDS reads of one step
// are in flight together instead of one dependent chain per query.
template <int ITEMS, typename T>
__device__ __forceinline__ void upper_entries(const T *off, int n, const T (&s)[ITEMS], int (&e)[ITEMS]) {
    int hi[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        e[j] = 0;
        hi[j] = n - 1;
    }
    const int steps = n <= 1 ? 0 : 32 - __clz(n - 1);
    for (int st = 0; st < steps; ++st) {
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const int mid = (e[j] + hi[j] + 1) >> 1;
            if (off[mid] <= s[j]) e[j] = mid;
            else hi[j] = mid - 1;
        }
    }
}

// Entries of products dealt to a wave in groups of 64 consecutive products:
// lane l holds product g0 + j * stride + l of group j (g0 and stride uniform
// across the wave, which calls this with every lane; off[0] <= g0, off
// nondecreasing, n entries).  Lane j finds the entry holding group j's first
// product by a binary search (one search per lane, not one per item), and
// every entry starting inside the group marks its first lane in `mark` (this
// wave's 64 ints; among entries starting at one product, empty ones first,
// the largest index wins): a prefix maximum over the lanes gives each lane
// its entry.  One LDS read per lane and group instead of a search per item.
template <int ITEMS, int MARKS = 1, typename T>
__device__ __forceinline__ void group_entries(const T *off, int n, T g0, T stride, int *mark, int (&e)[ITEMS]) {
    static_assert(ITEMS % MARKS == 0, "groups are marked MARKS at a time");
    const int lane = lane_id();
    int first;
    {
        const T q = g0 + (T)min(lane, ITEMS - 1) * stride;
        T qa[1] = {q};
        int ea[1];
        upper_entries<1>(off, n, qa, ea);
        first = ea[0];
    }
    // MARKS groups at a time, each in its own 64 marks (mark holds MARKS * 64
    // ints), so that their LDS round trips overlap.
#pragma unroll
    for (int j0 = 0; j0 < ITEMS; j0 += MARKS) {
#pragma unroll
        for (int m = 0; m < MARKS; ++m) mark[m * 64 + lane] = -1;
        wave_sync();
#pragma unroll
        for (int m = 0; m < MARKS; ++m) {
            const int j = j0 + m;
            const int e0 = __builtin_amdgcn_readlane(first, j);
            const T g = g0 + (T)j * stride;
            const T lim = g + 64;
            for (int c0 = e0 + 1; c0 < n; c0 += 64) {
                const int c = c0 + lane;
                const T o = c < n ? off[c] : lim;
                if (o < lim) atomicMax(&mark[m * 64 + (int)(o - g)], c);
                if (!((__ballot(o < lim) >> 63) & 1ull)) break;  // lane 63's entry starts past the group: no later one starts inside
            }
        }
        wave_sync();
#pragma unroll
        for (int m = 0; m < MARKS; ++m)
            e[j0 + m] = max(__builtin_amdgcn_readlane(first, j0 + m), wave_incl_max_dpp(mark[m * 64 + lane]));
        wave_sync();  // every lane has read the marks before they are reset
    }
}

// Where the products of a row go, slot by slot: slot j of thread tid holds the
// product (A entry e[j], B offset t[j]) when ok[j].  Entry-major when every
// referenced B row fits a power-of-two lane group (L lanes per A entry, no
// search, contiguous B reads); otherwise products are dealt out in serial
// order (s = tid + j*TPB) and located by a binary search over the offsets.
// off[e] is the exclusive offset of entry e's products, off[na] = P.
template <int TPB, int ITEMS>
__device__ __forceinline__ void slot_products(int na, int P, int L, const int *off, int (&e)[ITEMS], int (&t)[ITEMS],
                                              bool (&ok)[ITEMS]) {
    const int tid = threadIdx.x;
    int Lp = 1;
    while (Lp < L) Lp <<= 1;
    Lp = uni(Lp);
    const int logL = 31 - __clz(Lp);
    const int G = Lp <= TPB ? (TPB >> logL) : 1;
    const bool entry_major = Lp <= TPB && (na + G - 1) / G <= ITEMS;
    if (entry_major) {
        const int sub = tid & (Lp - 1);
        const int grp = tid >> logL;
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const int ee = j * G + grp;
            const bool in = ee < na;
            const int len = in ? off[ee + 1] - off[ee] : 0;
            ok[j] = in && sub < len;
            e[j] = in ? ee : 0;
            t[j] = sub;
        }
    } else {
        int sx[ITEMS];
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            sx[j] = tid + j * TPB;
            ok[j] = sx[j] < P;
        }
        upper_entries<ITEMS>(off, na, sx, e);
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            if (!ok[j]) e[j] = 0;
            t[j] = ok[j] ? sx[j] - off[e[j]] : 0;
        }
    }
}

__device__ __forceinline__ int classify(int64_t ub, int64_t na) {
    if (ub == 0) return kBinEmpty;
    if (ub <= kCapW1 && na <= kCapW1) return kBinW1;
    if (ub <= kCapW2 && na <= kCapW2) return kBinW2;
    if (ub <= kCapW3 && na <= kCapW3) return kBinW3;
    if (ub <= kCapW4 && na <= kCapW4) return kBinW4;
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

// Dynamic row assignment for the large-row kernels: the workgroup takes the
// next list entry from a counter in the control block (zeroed with it).  Rows
// differ by orders of magnitude in work, and the heaviest come first in row
// order on power-law matrices, so taking rows on demand balances the tail.
// No workgroup ever waits on another.
__device__ __forceinline__ int next_large_row(int *counter, int *s_slot) {
    if (threadIdx.x == 0) *s_slot = atomicAdd(counter, 1);
    __syncthreads();
    const int i = *s_slot;
    __syncthreads();  // every thread has read the slot before it is reused
    return i;
}

// Rows of one bin: list and count (filled by the analyze pass).
__device__ __forceinline__ const RowRec *bin_list(const Workspace &ws, int bin) {
    return ws.bin_recs + (size_t)bin * (size_t)ws.list_stride;
}
// The large rows heaviest first (bin_count[kBinLarge] records), ordered by
// the analyze reduction into the storage of the empty bin.
__device__ __forceinline__ const RowRec *large_list(const Workspace &ws) { return bin_list(ws, kBinEmpty); }

// Record i (block-uniform i) is fetched as one dword per lane (lanes 0..3 of
// each wave), i.e. with a vector load whose latency is tracked apart from LDS
// traffic, and decoded with shuffles only when it is needed.
__device__ __forceinline__ int rec_word(const RowRec *list, int i) {
    return reinterpret_cast<const int *>(list + i)[threadIdx.x & 3];
}
__device__ __forceinline__ RowRec rec_decode(int w) {
    RowRec r;
    r.row = __builtin_amdgcn_readlane(w, 0);
    r.na = __builtin_amdgcn_readlane(w, 1);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane(w, 2);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane(w, 3);
    r.a0 = (int64_t)(((uint64_t)hi << 32) | lo);
    return r;
}

// Start of the B row A entry e references (the analyze pass's stream), read
// as 32-bit starts when NARROW (mode bit kWsNarrow).  Hot kernels are
// instantiated for both formats, so they never branch on it; cold paths use
// entry_bstart_any.
template <bool NARROW>
__device__ __forceinline__ int64_t entry_bstart(const Workspace &ws, int64_t e) {
    if constexpr (NARROW) return (int64_t)__builtin_nontemporal_load(reinterpret_cast<const uint32_t *>(ws.ebst) + e);
    else return __builtin_nontemporal_load(&ws.ebst[e]);
}
__device__ __forceinline__ int64_t entry_bstart_any(const Workspace &ws, int64_t e) {
    return (ws.mode & kWsNarrow) ? entry_bstart<true>(ws, e) : entry_bstart<false>(ws, e);
}

// A row's per-entry metadata held in registers (thread e holds entry e), so
// the next row's metadata can be in flight while the current row computes.
struct MetaPre {
    int64_t bst;
    int len;
    double av;
};

// Loads are unconditional (lanes past the last entry re-read the row's last
// entry) so the destination registers never need a prior initialisation,
// which would force a wait on loads still in flight.
template <bool VALUES, bool NARROW>
__device__ __forceinline__ MetaPre meta_prefetch(const Operands &op, const Workspace &ws, const RowRec &r) {
    MetaPre m;
    const int64_t e = r.a0 + min((int)threadIdx.x, max(r.na - 1, 0));
    m.bst = entry_bstart<NARROW>(ws, e);
    m.len = __builtin_nontemporal_load(&ws.elen[e]);
    m.av = VALUES ? __builtin_nontemporal_load(&op.aval[e]) : 0.0;
    return m;
}

// Row metadata into LDS: B starts, (A values), exclusive product offsets
// off[0..na] and the longest referenced B row.  Entries < TPB come from the
// prefetched registers when `pre_ok`.
template <int TPB, bool VALUES, bool NARROW>
__device__ __forceinline__ void meta_to_lds(const Operands &op, const Workspace &ws, const RowRec &r,
                                            const MetaPre &pre, bool pre_ok, int64_t *s_bst, double *s_av,
                                            int *s_off, int *s_wsi, int &P, int &L) {
    const int tid = threadIdx.x;
    P = 0;
    L = 0;
    for (int e0 = 0; e0 < r.na; e0 += TPB) {
        const int e = e0 + tid;
        int len = 0;
        if (e < r.na) {
            int64_t bst;
            double av = 0.0;
            if (pre_ok && e0 == 0) {
                bst = pre.bst;
                len = pre.len;
                av = pre.av;
            } else {
                bst = entry_bstart<NARROW>(ws, r.a0 + e);
                len = __builtin_nontemporal_load(&ws.elen[r.a0 + e]);
                if (VALUES) av = __builtin_nontemporal_load(&op.aval[r.a0 + e]);
            }
            s_bst[e] = bst;
            if (VALUES) s_av[e] = av;
            L = max(L, len);
        }
        int tot;
        const int ex = block_excl_scan<TPB>(len, s_wsi, tot);
        if (e < r.na) s_off[e] = P + ex;
        P += tot;
    }
    P = uni(P);
    L = uni(block_max<TPB>(L, s_wsi));
}

template <typename K>
int occupancy_grid(K kernel, int tpb, int cus, long long want, size_t dyn_lds = 0) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, tpb, dyn_lds) != hipSuccess || per_cu < 1)
        per_cu = 1;
    long long g = (long long)per_cu * cus;
    if (want > 0 && want < g) g = want;
    return (int)(g < 1 ? 1 : g);
}

}  // namespace dev
}  // namespace spg
