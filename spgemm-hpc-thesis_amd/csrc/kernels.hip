// kernels.hip -- gfx950 kernels of the local SpGEMM product.
//
// Semantics (malchioman/SpGEMM-hpc-thesis, src/baselines/spgemm_exchange.cpp:274-309):
// for each output row, products a*b are formed for A entries in stored order and,
// within each, B entries in stored order; a column's first product initialises
// its sum and later ones are added left to right; columns come out ascending and
// sums equal to 0.0 are dropped.  Bit-exactness therefore needs every output
// coordinate summed sequentially in that serial order with separately rounded
// multiply and add (this file is built with -ffp-contract=off), so there are no
// value atomics and no tree reductions anywhere.
//
// Per-row method ("ordered ESC in LDS"): the row's products are expanded in
// serial order (product index s), bucketed by column range with LDS counters,
// scattered, and each thread then owns a contiguous run of buckets which it
// insertion-sorts by (column, s) and folds left to right.  The sort key carries
// s, so the fold order equals the serial order whatever order the LDS atomics
// placed items in.  Rows too large for one LDS pass are processed in chunks of
// the serial order; each chunk is sorted the same way and folded onto a dense
// per-workgroup accumulator in HBM (columns windowed), so the running sum of a
// column continues exactly where the previous chunk left it.
#include "internal.hpp"

#include <climits>

namespace spg {
namespace {

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

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

// Exclusive scan across a TPB-thread block.  `ws` holds TPB/64 elements.
// Contains barriers when TPB > 64; must be reached by every thread.
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
    if (ub <= kCapWave1 && na <= kCapWave1) return kBinWave1;
    if (ub <= kCapWave2 && na <= kCapWave2) return kBinWave2;
    if (ub <= kCapBlock && na <= kCapBlock) return kBinBlock;
    return kBinLarge;
}

// Presence bitmaps are updated with L2 atomics by many threads; plain loads
// could hit a stale L1 line, so reads and clears go to L2 as well.
__device__ __forceinline__ uint32_t ld_bits(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_bits(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ unsigned long long lanemask_lt() {
    return (1ull << lane_id()) - 1ull;
}

// ---------------------------------------------------------------------------
// Control-block reset (a kernel rather than a memset so the sequence can be
// captured in a graph).
// ---------------------------------------------------------------------------
__global__ void k_zero_control(Control *c) {
    int *p = reinterpret_cast<int *>(c);
    for (int i = threadIdx.x; i < (int)(sizeof(Control) / sizeof(int)); i += blockDim.x) p[i] = 0;
}

__global__ void k_zero_words(uint32_t *p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

// ---------------------------------------------------------------------------
// Analyze: per-row upper bound P_row (scalarMultiplicationCount per row,
// spgemm_common.cpp:474-487), A-entry count, bin; wave-aggregated counters.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_analyze(Operands op, Workspace ws) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    int b = -1;
    int64_t u = 0;
    unsigned flags = 0;
    if (r < op.arows) {
        const int64_t a0 = op.aptr[r], a1 = op.aptr[r + 1];
        for (int64_t e = a0; e < a1; ++e) {
            int64_t s, en;
            brow(op, op.acol[e], s, en, flags);
            u += en - s;
        }
        b = classify(u, a1 - a0);
        ws.ub[r] = u;
        ws.bin[r] = (uint8_t)b;
        if (b == kBinEmpty) {
            ws.snnz[r] = 0;
            ws.annz[r] = 0;
        }
    }
    const unsigned long long tot = (unsigned long long)wave_sum(u);
    unsigned long long fl = flags;
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) fl |= __shfl_xor(fl, d, 64);
    if (lane_id() == 0) {
        if (tot) atomicAdd(&ws.ctrl->products, tot);
        if (fl) atomicOr(&ws.ctrl->flags, fl);
    }
#pragma unroll
    for (int bb = 0; bb < kNumBins; ++bb) {
        const unsigned long long m = __ballot(b == bb);
        if (m && lane_id() == 0) atomicAdd(&ws.ctrl->bin_count[bb], (int)__popcll(m));
    }
}

// Offsets of each bin's row list, then the lists (empty rows are not listed).
__global__ __launch_bounds__(256) void k_bin_fill(Operands op, Workspace ws) {
    __shared__ int s_off[kNumBins + 1];
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int bb = 0; bb < kNumBins; ++bb) {
            s_off[bb] = acc;
            if (bb != kBinEmpty) acc += ws.ctrl->bin_count[bb];
        }
        s_off[kNumBins] = acc;
        if (blockIdx.x == 0)
            for (int bb = 0; bb <= kNumBins; ++bb) ws.ctrl->bin_offset[bb] = s_off[bb];
    }
    __syncthreads();
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = r < op.arows ? (int)ws.bin[r] : -1;
#pragma unroll
    for (int bb = 1; bb < kNumBins; ++bb) {
        const unsigned long long m = __ballot(b == bb);
        if (!m) continue;
        const int leader = __ffsll((long long)m) - 1;
        int base = 0;
        if (lane_id() == leader) base = atomicAdd(&ws.ctrl->bin_cursor[bb], (int)__popcll(m));
        base = __shfl(base, leader, 64);
        if (b == bb) ws.bin_rows[s_off[bb] + base + (int)__popcll(m & lanemask_lt())] = r;
    }
}

// ---------------------------------------------------------------------------
// One row per TPB-thread workgroup, all of the row's products in LDS.
// NUM=false: symbolic (distinct columns -> snnz).  NUM=true: numeric.
// LDS reuse: during expansion s_key holds each A entry's B start, s_val its A
// value and s_cnt the product offsets; afterwards they hold the items.
// ---------------------------------------------------------------------------
template <int TPB, int CAP, bool NUM>
__global__ __launch_bounds__(TPB) void k_row_esc(Operands op, Workspace ws, int binid, const int64_t *crowptr,
                                                 int32_t *ccol, double *cval) {
    constexpr int ITEMS = CAP / TPB;
    static_assert(CAP % TPB == 0, "CAP must be a multiple of TPB");
    __shared__ uint64_t s_key[CAP];
    __shared__ double s_val[NUM ? CAP : 1];
    __shared__ int s_cnt[CAP + 1];
    __shared__ int s_wsi[2 * (TPB / 64) + 2];
    int64_t *s_bst = reinterpret_cast<int64_t *>(s_key);
    const int tid = threadIdx.x;
    const int beg = ws.ctrl->bin_offset[binid];
    const int end = beg + ws.ctrl->bin_count[binid];
    unsigned long long stored = 0;
    unsigned flags = 0;

    for (int i = beg + (int)blockIdx.x; i < end; i += (int)gridDim.x) {
        const int row = ws.bin_rows[i];
        const int64_t a0 = op.aptr[row];
        const int na = (int)(op.aptr[row + 1] - a0);  // <= CAP by binning

        // --- A-entry metadata and the exclusive scan of B row lengths.
        const int ipt = (na + TPB - 1) / TPB;
        const int e0 = tid * ipt;
        int mylen = 0;
        for (int j = 0; j < ipt; ++j) {
            const int e = e0 + j;
            if (e < na) {
                int64_t s, en;
                brow(op, op.acol[a0 + e], s, en, flags);
                const int len = (int)(en - s);
                s_bst[e] = s;
                if constexpr (NUM) s_val[e] = op.aval[a0 + e];
                s_cnt[e] = len;
                mylen += len;
            }
        }
        int P;
        int base = block_excl_scan<TPB>(mylen, s_wsi, P);
        for (int j = 0; j < ipt; ++j) {
            const int e = e0 + j;
            if (e < na) {
                const int l = s_cnt[e];
                s_cnt[e] = base;
                base += l;
            }
        }
        if (P > CAP) {  // cannot happen: binning used the same lengths
            flags |= kErrInternal;
            P = CAP;
        }
        __syncthreads();

        // --- Expansion in serial order: item s = tid + j*TPB (coalesced B reads).
        int icol[ITEMS];
        double ival[ITEMS];
        int cmn = INT_MAX, cmx = INT_MIN;
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const int s = tid + j * TPB;
            icol[j] = 0;
            ival[j] = 0.0;
            if (s < P) {
                const int e = upper_entry(s_cnt, na, s);
                const int64_t idx = s_bst[e] + (s - s_cnt[e]);
                const int c = op.bcol[idx];
                icol[j] = c;
                if constexpr (NUM) ival[j] = s_val[e] * op.bval[idx];
                cmn = min(cmn, c);
                cmx = max(cmx, c);
            }
        }
        block_minmax<TPB>(cmn, cmx, s_wsi);
        __syncthreads();  // metadata no longer needed

        // --- Bucket by column range (power-of-two buckets, nb in [TPB, CAP]).
        int nb = TPB;
        while (nb < P && nb < CAP) nb <<= 1;
        const int lognb = 31 - __clz(nb);
        const uint64_t range = (uint64_t)((int64_t)cmx - (int64_t)cmn) + 1ull;
        const int lr = range <= 1 ? 0 : 64 - __clzll((long long)(range - 1));
        const int sh = lr > lognb ? lr - lognb : 0;
        for (int t = tid; t < nb; t += TPB) s_cnt[t] = 0;
        __syncthreads();
        int bk[ITEMS], pos[ITEMS];
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const int s = tid + j * TPB;
            bk[j] = 0;
            pos[j] = 0;
            if (s < P) {
                bk[j] = (int)((uint32_t)((int64_t)icol[j] - (int64_t)cmn) >> sh);
                pos[j] = atomicAdd(&s_cnt[bk[j]], 1);
            }
        }
        __syncthreads();
        const int bpt = nb / TPB;
        int loc = 0;
        for (int q = 0; q < bpt; ++q) loc += s_cnt[tid * bpt + q];
        int tot;
        int run = block_excl_scan<TPB>(loc, s_wsi, tot);
        for (int q = 0; q < bpt; ++q) {
            const int c = s_cnt[tid * bpt + q];
            s_cnt[tid * bpt + q] = run;
            run += c;
        }
        if (tid == 0) s_cnt[nb] = P;
        __syncthreads();
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const int s = tid + j * TPB;
            if (s < P) {
                const int d = s_cnt[bk[j]] + pos[j];
                s_key[d] = ((uint64_t)(uint32_t)icol[j] << 32) | (uint32_t)s;
                if constexpr (NUM) s_val[d] = ival[j];
            }
        }
        __syncthreads();

        // --- Own run of buckets: insertion sort by (column, s), ordered fold.
        const int lo = s_cnt[tid * bpt];
        const int hi = s_cnt[tid * bpt + bpt];
        for (int x = lo + 1; x < hi; ++x) {
            const uint64_t k = s_key[x];
            double v = 0.0;
            if constexpr (NUM) v = s_val[x];
            int y = x - 1;
            while (y >= lo && s_key[y] > k) {
                s_key[y + 1] = s_key[y];
                if constexpr (NUM) s_val[y + 1] = s_val[y];
                --y;
            }
            s_key[y + 1] = k;
            if constexpr (NUM) s_val[y + 1] = v;
        }
        int uniq = 0, nz = 0, w = lo;
        for (int x = lo; x < hi;) {
            const uint32_t c = (uint32_t)(s_key[x] >> 32);
            double v = 0.0;
            if constexpr (NUM) v = s_val[x];
            ++x;
            while (x < hi && (uint32_t)(s_key[x] >> 32) == c) {
                if constexpr (NUM) v = v + s_val[x];
                ++x;
            }
            ++uniq;
            if constexpr (NUM) {
                if (v != 0.0) {
                    s_key[w] = c;
                    s_val[w] = v;
                    ++w;
                    ++nz;
                }
            }
        }
        if constexpr (!NUM) {
            const int total = block_sum<TPB>(uniq, s_wsi);
            if (tid == 0) ws.snnz[row] = total;
        } else {
            int total;
            const int off = block_excl_scan<TPB>(nz, s_wsi, total);
            const int64_t cbase = crowptr[row];
            const int64_t lim = crowptr[row + 1] - cbase;
            for (int q = lo; q < w; ++q) {
                const int64_t o = off + (q - lo);
                if (o < lim) {
                    ccol[cbase + o] = (int32_t)(uint32_t)s_key[q];
                    cval[cbase + o] = s_val[q];
                }
            }
            if (tid == 0) {
                ws.annz[row] = total;
                stored += (unsigned long long)total;
                if (total > lim) flags |= kErrInternal;
            }
        }
        __syncthreads();  // LDS is reused by the next row
    }
    if (NUM && tid == 0 && stored) atomicAdd(&ws.ctrl->actual, stored);
    if (flags) atomicOr(&ws.ctrl->flags, (unsigned long long)flags);
}

// ---------------------------------------------------------------------------
// Large rows.  A row's A entries are taken in batches of up to CAP entries
// (product offsets kept in LDS), each batch's products in chunks of up to CAP
// in serial order.  Columns are processed in windows of ws.lwidth.
// ---------------------------------------------------------------------------
template <int TPB, int CAP>
__device__ __forceinline__ int64_t large_batch_meta(const Operands &op, int64_t e0, int nbt, int64_t *s_bst,
                                                    double *s_av, int64_t *s_off, int64_t *s_ws64,
                                                    unsigned &flags, bool want_values) {
    const int tid = threadIdx.x;
    const int ipt = (nbt + TPB - 1) / TPB;
    const int f = tid * ipt;
    int64_t mylen = 0;
    for (int j = 0; j < ipt; ++j) {
        const int e = f + j;
        if (e < nbt) {
            int64_t s, en;
            brow(op, op.acol[e0 + e], s, en, flags);
            s_bst[e] = s;
            if (want_values) s_av[e] = op.aval[e0 + e];
            s_off[e] = en - s;
            mylen += en - s;
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
    const int beg = ws.ctrl->bin_offset[kBinLarge];
    const int end = beg + ws.ctrl->bin_count[kBinLarge];
    const int64_t W = ws.lwidth;
    uint32_t *bits = ws.lbits + (size_t)blockIdx.x * (size_t)(W / 32);
    unsigned flags = 0;
    for (int i = beg + (int)blockIdx.x; i < end; i += (int)gridDim.x) {
        const int row = ws.bin_rows[i];
        const int64_t a0 = op.aptr[row], a1 = op.aptr[row + 1];
        int64_t cnt = 0;
        for (int64_t w0 = 0; w0 < op.bcols; w0 += W) {
            const int64_t wend = min(w0 + W, (int64_t)op.bcols);
            for (int64_t e0 = a0; e0 < a1; e0 += CAP) {
                const int nbt = (int)min((int64_t)CAP, a1 - e0);
                const int64_t Pb = large_batch_meta<TPB, CAP>(op, e0, nbt, s_bst, nullptr, s_off, s_ws64, flags, false);
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
__global__ __launch_bounds__(TPB) void k_large_numeric(Operands op, Workspace ws, const int64_t *crowptr,
                                                       int32_t *ccol, double *cval) {
    constexpr int ITEMS = CAP / TPB;
    __shared__ int64_t s_bst[CAP];
    __shared__ double s_av[CAP];
    __shared__ int64_t s_off[CAP + 1];
    __shared__ uint64_t s_key[CAP];
    __shared__ double s_val[CAP];
    __shared__ int s_cnt[CAP + 1];
    __shared__ int s_wsi[2 * (TPB / 64) + 2];
    __shared__ int64_t s_ws64[TPB / 64 + 1];
    const int tid = threadIdx.x;
    const int beg = ws.ctrl->bin_offset[kBinLarge];
    const int end = beg + ws.ctrl->bin_count[kBinLarge];
    const int64_t W = ws.lwidth;
    uint32_t *bits = ws.lbits + (size_t)blockIdx.x * (size_t)(W / 32);
    double *acc = ws.lacc + (size_t)blockIdx.x * (size_t)W;
    unsigned flags = 0;
    unsigned long long stored = 0;
    for (int i = beg + (int)blockIdx.x; i < end; i += (int)gridDim.x) {
        const int row = ws.bin_rows[i];
        const int64_t a0 = op.aptr[row], a1 = op.aptr[row + 1];
        const int64_t cbase = crowptr[row];
        const int64_t lim = crowptr[row + 1] - cbase;
        int64_t outn = 0;
        for (int64_t w0 = 0; w0 < op.bcols; w0 += W) {
            const int64_t wend = min(w0 + W, (int64_t)op.bcols);
            for (int64_t e0 = a0; e0 < a1; e0 += CAP) {
                const int nbt = (int)min((int64_t)CAP, a1 - e0);
                const int64_t Pb = large_batch_meta<TPB, CAP>(op, e0, nbt, s_bst, s_av, s_off, s_ws64, flags, true);
                for (int64_t c0 = 0; c0 < Pb; c0 += CAP) {
                    const int Pc = (int)min((int64_t)CAP, Pb - c0);
                    int icol[ITEMS];
                    double ival[ITEMS];
                    bool ok[ITEMS];
                    int cmn = INT_MAX, cmx = INT_MIN, nvalid = 0;
#pragma unroll
                    for (int j = 0; j < ITEMS; ++j) {
                        const int s = tid + j * TPB;
                        icol[j] = 0;
                        ival[j] = 0.0;
                        ok[j] = false;
                        if (s < Pc) {
                            const int64_t g = c0 + s;
                            const int e = upper_entry(s_off, nbt, g);
                            const int64_t idx = s_bst[e] + (g - s_off[e]);
                            const int c = op.bcol[idx];
                            if (c >= w0 && c < wend) {
                                icol[j] = c;
                                ival[j] = s_av[e] * op.bval[idx];
                                ok[j] = true;
                                cmn = min(cmn, c);
                                cmx = max(cmx, c);
                                ++nvalid;
                            }
                        }
                    }
                    const int V = block_sum<TPB>(nvalid, s_wsi);
                    if (V == 0) continue;  // uniform across the block
                    block_minmax<TPB>(cmn, cmx, s_wsi);
                    int nb = TPB;
                    while (nb < V && nb < CAP) nb <<= 1;
                    const int lognb = 31 - __clz(nb);
                    const uint64_t range = (uint64_t)((int64_t)cmx - (int64_t)cmn) + 1ull;
                    const int lr = range <= 1 ? 0 : 64 - __clzll((long long)(range - 1));
                    const int sh = lr > lognb ? lr - lognb : 0;
                    for (int t = tid; t < nb; t += TPB) s_cnt[t] = 0;
                    __syncthreads();
                    int bk[ITEMS], pos[ITEMS];
#pragma unroll
                    for (int j = 0; j < ITEMS; ++j) {
                        bk[j] = 0;
                        pos[j] = 0;
                        if (ok[j]) {
                            bk[j] = (int)((uint32_t)((int64_t)icol[j] - (int64_t)cmn) >> sh);
                            pos[j] = atomicAdd(&s_cnt[bk[j]], 1);
                        }
                    }
                    __syncthreads();
                    const int bpt = nb / TPB;
                    int loc = 0;
                    for (int q = 0; q < bpt; ++q) loc += s_cnt[tid * bpt + q];
                    int tot;
                    int run = block_excl_scan<TPB>(loc, s_wsi, tot);
                    for (int q = 0; q < bpt; ++q) {
                        const int c = s_cnt[tid * bpt + q];
                        s_cnt[tid * bpt + q] = run;
                        run += c;
                    }
                    if (tid == 0) s_cnt[nb] = V;
                    __syncthreads();
#pragma unroll
                    for (int j = 0; j < ITEMS; ++j) {
                        if (ok[j]) {
                            const int s = tid + j * TPB;
                            const int d = s_cnt[bk[j]] + pos[j];
                            s_key[d] = ((uint64_t)(uint32_t)icol[j] << 32) | (uint32_t)s;
                            s_val[d] = ival[j];
                        }
                    }
                    __syncthreads();
                    const int lo = s_cnt[tid * bpt];
                    const int hi = s_cnt[tid * bpt + bpt];
                    for (int x = lo + 1; x < hi; ++x) {
                        const uint64_t k = s_key[x];
                        const double v = s_val[x];
                        int y = x - 1;
                        while (y >= lo && s_key[y] > k) {
                            s_key[y + 1] = s_key[y];
                            s_val[y + 1] = s_val[y];
                            --y;
                        }
                        s_key[y + 1] = k;
                        s_val[y + 1] = v;
                    }
                    for (int x = lo; x < hi;) {
                        const uint32_t c = (uint32_t)(s_key[x] >> 32);
                        const uint32_t rel = (uint32_t)((int64_t)(int32_t)c - w0);
                        const uint32_t bit = 1u << (rel & 31u);
                        const bool present = (ld_bits(&bits[rel >> 5]) & bit) != 0u;
                        double v = present ? acc[rel] : 0.0;
                        while (x < hi && (uint32_t)(s_key[x] >> 32) == c) {
                            v = v + s_val[x];
                            ++x;
                        }
                        acc[rel] = v;
                        if (!present) atomicOr(&bits[rel >> 5], bit);
                    }
                    __syncthreads();  // chunk folded; LDS and accumulator state consistent
                }
            }
            // Extraction of the window in ascending column order, zeros dropped.
            const int64_t nwords = (wend - w0 + 31) / 32;
            const int64_t wpt = (nwords + TPB - 1) / TPB;
            const int64_t q0 = (int64_t)tid * wpt;
            const int64_t q1 = min(q0 + wpt, nwords);
            int64_t mine = 0;
            for (int64_t q = q0; q < q1; ++q) {
                uint32_t m = ld_bits(&bits[q]);
                while (m) {
                    const int b = __ffs((int)m) - 1;
                    m &= m - 1u;
                    if (acc[q * 32 + b] != 0.0) ++mine;
                }
            }
            int64_t wtotal;
            int64_t o = outn + block_excl_scan<TPB>(mine, s_ws64, wtotal);
            for (int64_t q = q0; q < q1; ++q) {
                uint32_t m = ld_bits(&bits[q]);
                while (m) {
                    const int b = __ffs((int)m) - 1;
                    m &= m - 1u;
                    const double v = acc[q * 32 + b];
                    if (v != 0.0) {
                        if (o < lim) {
                            ccol[cbase + o] = (int32_t)(w0 + q * 32 + b);
                            cval[cbase + o] = v;
                        }
                        ++o;
                    }
                }
                st_bits(&bits[q], 0u);
            }
            outn += wtotal;
            __threadfence();
            __syncthreads();
        }
        if (tid == 0) {
            ws.annz[row] = outn;
            stored += (unsigned long long)outn;
            if (outn > lim) flags |= kErrInternal;
        }
    }
    if (tid == 0 && stored) atomicAdd(&ws.ctrl->actual, stored);
    if (flags) atomicOr(&ws.ctrl->flags, (unsigned long long)flags);
}

// ---------------------------------------------------------------------------
// Exclusive scan of int64 counts (three passes, 2048 elements per block).
// ---------------------------------------------------------------------------
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

// ---------------------------------------------------------------------------
// Compaction (only when exact zeros were dropped): one wave per row.
// ---------------------------------------------------------------------------
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

template <typename K>
int occupancy_grid(K kernel, int tpb, int cus, int want_rows) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, tpb, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    long long g = (long long)per_cu * cus;
    if (want_rows < g) g = want_rows;
    return (int)(g < 1 ? 1 : g);
}

}  // namespace

void launch_zero_control(Control *ctrl, hipStream_t stream) {
    hipLaunchKernelGGL(k_zero_control, dim3(1), dim3(64), 0, stream, ctrl);
}

void launch_zero_bytes(void *p, size_t bytes, hipStream_t stream) {
    const size_t n = bytes / 4;
    if (!n) return;
    size_t blocks = (n + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    hipLaunchKernelGGL(k_zero_words, dim3((unsigned)blocks), dim3(256), 0, stream, static_cast<uint32_t *>(p), n);
}

void launch_analyze(const Operands &op, const Workspace &ws, hipStream_t stream) {
    if (op.arows <= 0) return;
    const int blocks = (op.arows + 255) / 256;
    hipLaunchKernelGGL(k_analyze, dim3(blocks), dim3(256), 0, stream, op, ws);
}

void launch_bin_fill(const Operands &op, const Workspace &ws, hipStream_t stream) {
    const int blocks = op.arows > 0 ? (op.arows + 255) / 256 : 1;
    hipLaunchKernelGGL(k_bin_fill, dim3(blocks), dim3(256), 0, stream, op, ws);
}

void launch_symbolic(const Operands &op, const Workspace &ws, const int *bin_counts, int cus, hipStream_t stream) {
    if (bin_counts[kBinWave1] > 0) {
        auto k = k_row_esc<64, kCapWave1, false>;
        const int g = occupancy_grid(k, 64, cus, bin_counts[kBinWave1]);
        hipLaunchKernelGGL(k, dim3(g), dim3(64), 0, stream, op, ws, (int)kBinWave1, nullptr, nullptr, nullptr);
    }
    if (bin_counts[kBinWave2] > 0) {
        auto k = k_row_esc<64, kCapWave2, false>;
        const int g = occupancy_grid(k, 64, cus, bin_counts[kBinWave2]);
        hipLaunchKernelGGL(k, dim3(g), dim3(64), 0, stream, op, ws, (int)kBinWave2, nullptr, nullptr, nullptr);
    }
    if (bin_counts[kBinBlock] > 0) {
        auto k = k_row_esc<256, kCapBlock, false>;
        const int g = occupancy_grid(k, 256, cus, bin_counts[kBinBlock]);
        hipLaunchKernelGGL(k, dim3(g), dim3(256), 0, stream, op, ws, (int)kBinBlock, nullptr, nullptr, nullptr);
    }
    if (bin_counts[kBinLarge] > 0) {
        const int g = bin_counts[kBinLarge] < ws.lblocks ? bin_counts[kBinLarge] : ws.lblocks;
        hipLaunchKernelGGL((k_large_symbolic<kTpbLarge, kCapLarge>), dim3(g), dim3(kTpbLarge), 0, stream, op, ws);
    }
}

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

void launch_numeric(const Operands &op, const Workspace &ws, const int *bin_counts, int cus,
                    const int64_t *crowptr, int32_t *ccol, double *cval, hipStream_t stream) {
    if (bin_counts[kBinWave1] > 0) {
        auto k = k_row_esc<64, kCapWave1, true>;
        const int g = occupancy_grid(k, 64, cus, bin_counts[kBinWave1]);
        hipLaunchKernelGGL(k, dim3(g), dim3(64), 0, stream, op, ws, (int)kBinWave1, crowptr, ccol, cval);
    }
    if (bin_counts[kBinWave2] > 0) {
        auto k = k_row_esc<64, kCapWave2, true>;
        const int g = occupancy_grid(k, 64, cus, bin_counts[kBinWave2]);
        hipLaunchKernelGGL(k, dim3(g), dim3(64), 0, stream, op, ws, (int)kBinWave2, crowptr, ccol, cval);
    }
    if (bin_counts[kBinBlock] > 0) {
        auto k = k_row_esc<256, kCapBlock, true>;
        const int g = occupancy_grid(k, 256, cus, bin_counts[kBinBlock]);
        hipLaunchKernelGGL(k, dim3(g), dim3(256), 0, stream, op, ws, (int)kBinBlock, crowptr, ccol, cval);
    }
    if (bin_counts[kBinLarge] > 0) {
        const int g = bin_counts[kBinLarge] < ws.lblocks ? bin_counts[kBinLarge] : ws.lblocks;
        hipLaunchKernelGGL((k_large_numeric<kTpbLarge, kCapLarge>), dim3(g), dim3(kTpbLarge), 0, stream, op, ws,
                           crowptr, ccol, cval);
    }
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
