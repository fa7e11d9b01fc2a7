// numeric.hip -- values of C in the reference's serial summation order.
//
// Semantics (src/baselines/spgemm_exchange.cpp:274-309): for each output row,
// products a*b are formed for A entries in stored order and, within each, B
// entries in stored order; a column's first product initialises its sum and
// later ones are added left to right; columns come out ascending and sums equal
// to 0.0 are dropped.  Bit-exactness needs every coordinate summed sequentially
// in that order with separately rounded multiply and add (built with
// -ffp-contract=off), so there are no value atomics and no tree reductions.
//
// Rows that fit LDS ("ordered ESC"): products are expanded with their serial
// index s, bucketed by column range with LDS counters and scattered.  Each item
// then computes its rank inside its (small) bucket by the key (column, s) --
// its final sorted position -- so the order the LDS atomics placed items in
// never matters and no thread loops over more than one bucket.  After the
// items move to their positions, the first position of each column folds the
// column's products left to right (serial order), and a scan of the nonzero
// flags in position order gives coalesced output offsets.
//
// Larger rows are taken in chunks of the serial order; each chunk is sorted the
// same way and folded onto a dense per-workgroup accumulator in HBM (columns in
// windows), so a column's running sum continues where the previous chunk left it.
#include "device_common.hpp"

#include <atomic>

namespace spg {
namespace {

using namespace dev;

template <int TPB, int CAP>
__device__ __forceinline__ void esc_row(const Operands &op, const Workspace &ws, const RowRec &cur, int P, int L,
                                        const int64_t *crowptr, int32_t *ccol, double *cval, double *s_a,
                                        int64_t *s_b64, int *s_c, int *s_wsi, uint64_t *s_ws64,
                                        unsigned long long &stored, unsigned &flags) {
    constexpr int ITEMS = CAP / TPB;
    // Bucketed items are keys (column - row minimum) << 32 | s: one LDS word
    // per item, ordered as (column, s), the bucket recomputed from the key.
    // Sorted columns then go to a dense int32 array over the dead keys.
    uint64_t *s_key = reinterpret_cast<uint64_t *>(s_b64);
    int32_t *s_col = reinterpret_cast<int32_t *>(s_b64);
    const int tid = threadIdx.x;
    const int na = cur.na;
    // Rows of more than 256 products load their output range with their
    // gathers, so its latency is not exposed after the sort; smaller rows load
    // it after the sort, where its registers cost no wave per SIMD.
    constexpr bool kEarlyRange = CAP >= 512;
    int64_t cbase = 0, lim = 0;
    if constexpr (kEarlyRange) {
        cbase = crowptr[cur.row];
        lim = crowptr[cur.row + 1] - cbase;
    }

    // --- expansion into registers: column, product, serial index s
    int icol[ITEMS], isx[ITEMS];
    double ival[ITEMS];
    int cmn = INT_MAX, cmx = INT_MIN;
    {
        // Every slot's LDS reads and B loads are issued before any product is
        // formed, so the row's gathers are in flight together.  Empty slots
        // read B entry 0 (the row's products exist, so B is not empty).
        int e[ITEMS], t[ITEMS];
        bool ok[ITEMS];
        slot_products<TPB, ITEMS>(na, P, L, s_c, e, t, ok);
        int64_t idx[ITEMS];
        double av[ITEMS], bv[ITEMS];
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            idx[j] = ok[j] ? s_b64[e[j]] + t[j] : 0;
            av[j] = s_a[e[j]];
            isx[j] = ok[j] ? s_c[e[j]] + t[j] : -1;
        }
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            icol[j] = op.bcol[idx[j]];
            bv[j] = op.bval[idx[j]];
        }
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            ival[j] = av[j] * bv[j];
            cmn = ok[j] ? min(cmn, icol[j]) : cmn;
            cmx = ok[j] ? max(cmx, icol[j]) : cmx;
        }
    }
    block_minmax<TPB>(cmn, cmx, s_wsi);
    __syncthreads();  // metadata no longer needed

    // --- buckets over the row's column range (power of two, nb in [TPB, CAP])
    int nb = TPB;
    while (nb < P && nb < CAP) nb <<= 1;
    nb = uni(nb);
    cmn = uni(cmn);
    cmx = uni(cmx);
    const int lognb = 31 - __clz(nb);
    const uint64_t range = (uint64_t)((int64_t)cmx - (int64_t)cmn) + 1ull;
    const int lr = range <= 1 ? 0 : 64 - __clzll((long long)(range - 1));
    const int sh = uni(lr > lognb ? lr - lognb : 0);
    for (int t = tid; t < nb; t += TPB) s_c[t] = 0;
    __syncthreads();
    int bk[ITEMS], pos[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        bk[j] = 0;
        pos[j] = 0;
        if (isx[j] >= 0) {
            bk[j] = (int)((uint32_t)((int64_t)icol[j] - (int64_t)cmn) >> sh);
            pos[j] = atomicAdd(&s_c[bk[j]], 1);
        }
    }
    __syncthreads();
    block_scan_runs<TPB>(s_c, nb, s_wsi);
    if (tid == 0) s_c[nb] = P;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        if (isx[j] >= 0) {
            const int d = s_c[bk[j]] + pos[j];
            s_key[d] = ((uint64_t)(uint32_t)((int64_t)icol[j] - (int64_t)cmn) << 32) | (uint32_t)isx[j];
            s_a[d] = ival[j];
        }
    }
    __syncthreads();

    // --- rank of each item inside its bucket by (column, s): its sorted position
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const int x = tid + j * TPB;
        isx[j] = -1;
        if (x < P) {
            const uint64_t k = s_key[x];
            const uint32_t rel = (uint32_t)(k >> 32);
            const int b = (int)(rel >> sh);
            const int st = s_c[b], en = s_c[b + 1];
            int rank = 0;
            for (int y = st; y < en; ++y) rank += s_key[y] < k;
            icol[j] = (int32_t)((int64_t)cmn + (int64_t)rel);
            ival[j] = s_a[x];
            isx[j] = st + rank;
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        if (isx[j] >= 0) {
            s_col[isx[j]] = icol[j];
            s_a[isx[j]] = ival[j];
        }
    }
    __syncthreads();

    // --- ordered fold at each column's first position, then the output: slot
    // j of thread tid is sorted position tid + j * TPB.  Every slot's column
    // and value are read first; a head folds its column's run in serial
    // order; one block scan of 16-bit per-slot flags (four slots per word)
    // gives every stored entry's place.
    if constexpr (!kEarlyRange) {
        cbase = crowptr[cur.row];
        lim = crowptr[cur.row + 1] - cbase;
    }
    int32_t cc[ITEMS];
    double vv[ITEMS];
    bool hd[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const int x = tid + j * TPB;
        const bool in = x < P;
        cc[j] = in ? s_col[x] : 0;
        vv[j] = in ? s_a[x] : 0.0;
        const int32_t pc = in && x > 0 ? s_col[x - 1] : 0;
        hd[j] = in && (x == 0 || pc != cc[j]);
    }
    uint64_t pk[(ITEMS + 3) / 4];
#pragma unroll
    for (int h = 0; h < (ITEMS + 3) / 4; ++h) pk[h] = 0;
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        if (hd[j]) {
            double v = vv[j];
            for (int y = tid + j * TPB + 1; y < P && s_col[y] == cc[j]; ++y) v = v + s_a[y];
            vv[j] = v;
        }
        const bool nz = hd[j] && (op.keep_zeros || vv[j] != 0.0);
        hd[j] = nz;
        pk[j / 4] |= (uint64_t)(nz ? 1u : 0u) << (16 * (j % 4));
    }
    uint64_t ex[(ITEMS + 3) / 4], tt[(ITEMS + 3) / 4];
#pragma unroll
    for (int h = 0; h < (ITEMS + 3) / 4; ++h) ex[h] = block_excl_scan<TPB>(pk[h], s_ws64, tt[h]);
    int64_t run = 0;
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        if (hd[j]) {
            const int64_t o = run + (int64_t)((ex[j / 4] >> (16 * (j % 4))) & 0xffffu);
            if (o < lim) {
                __builtin_nontemporal_store(cc[j], &ccol[cbase + o]);
                __builtin_nontemporal_store(vv[j], &cval[cbase + o]);
            }
        }
        run += (int64_t)((tt[j / 4] >> (16 * (j % 4))) & 0xffffu);
    }
    if (tid == 0) {
        ws.annz[cur.row] = run;
        stored += (unsigned long long)run;
        if (run > lim) flags |= kErrInternal;
    }
}

template <int TPB, int CAP>
struct EscLds {
    double a[CAP];     // meta: A value of entry e | items: product value
    int64_t b64[CAP];  // meta: B start of entry e | items: keys (column, s), then sorted columns
    int c[CAP + 1];    // meta: product offset of entry e | bucket counts, then bucket starts
    int wsi[2 * (TPB / 64) + 2];
    uint64_t ws64[TPB / 64];  // the output scan's per-wave totals
};

// One row of the software pipeline (see hash_step in symbolic.hip).
template <int TPB, int CAP, bool NARROW>
__device__ __forceinline__ void esc_step(const Operands &op, const Workspace &ws, const RowRec *list, int count,
                                         int G, int &i, RowRec &cur, const MetaPre &pcur, MetaPre &pnext,
                                         int &wnext, const int64_t *crowptr, int32_t *ccol, double *cval,
                                         EscLds<TPB, CAP> &sh, unsigned long long &stored, unsigned &flags) {
    const int tid = threadIdx.x;
    int P, L;
    meta_to_lds<TPB, true, NARROW>(op, ws, cur, pcur, cur.na <= TPB, sh.b64, sh.a, sh.c, sh.wsi, P, L);
    const RowRec nxt = i + G < count ? rec_decode(wnext) : cur;
    wnext = rec_word(list, min(i + 2 * G, count - 1));
    pnext = meta_prefetch<true, NARROW>(op, ws, nxt);
    if (P > CAP) flags |= kErrInternal;  // cannot happen: binning used the same lengths
    if (P <= CAP) {
        if (tid == 0) sh.c[cur.na] = P;
        __syncthreads();
        esc_row<TPB, CAP>(op, ws, cur, P, L, crowptr, ccol, cval, sh.a, sh.b64, sh.c, sh.wsi, sh.ws64, stored, flags);
    }
    __syncthreads();  // LDS is reused by the next row
    cur = nxt;
    i += G;
}

// Waves per SIMD the value kernels keep (1: no constraint): bin W4's wide
// format would otherwise take 193 VGPRs (two waves), bin W1's 93 (five).
template <int CAP>
constexpr int esc_min_waves() {
    return CAP == kCapW4 ? 3 : CAP == kCapW1 ? 6 : 1;
}

template <int TPB, int CAP, bool NARROW>
__global__ __launch_bounds__(TPB, esc_min_waves<CAP>()) void k_numeric_esc(Operands op, Workspace ws, int binid,
                                                                            const int64_t *crowptr, int32_t *ccol,
                                                                            double *cval) {
    static_assert(CAP % TPB == 0 && CAP <= 65536, "CAP must be a multiple of TPB and fit 16-bit indices");
    __shared__ EscLds<TPB, CAP> sh;
    const int tid = threadIdx.x;
    const RowRec *list = bin_list(ws, binid);
    const int count = ws.ctrl->bin_count[binid];
    const int G = (int)gridDim.x;
    unsigned long long stored = 0;
    unsigned flags = 0;
    int i = (int)blockIdx.x;
    if (i < count) {
        // Two-step unrolled loop with ping-pong prefetch registers (pa, pb).
        RowRec cur = rec_decode(rec_word(list, i));
        MetaPre pa = meta_prefetch<true, NARROW>(op, ws, cur), pb;
        int wnext = rec_word(list, min(i + G, count - 1));
        for (;;) {
            if (i >= count) break;
            esc_step<TPB, CAP, NARROW>(op, ws, list, count, G, i, cur, pa, pb, wnext, crowptr, ccol, cval, sh, stored,
                                       flags);
            if (i >= count) break;
            esc_step<TPB, CAP, NARROW>(op, ws, list, count, G, i, cur, pb, pa, wnext, crowptr, ccol, cval, sh, stored,
                                       flags);
        }
    }
    if (tid == 0 && stored) atomicAdd(&ws.ctrl->actual, stored);
    if (flags) atomicOr(&ws.ctrl->flags, (unsigned long long)flags);
}

// The large-row accumulator window belongs to one workgroup: a column's
// running value may have been written by another thread of the workgroup in
// an earlier chunk or row, with a workgroup barrier in between, so relaxed
// atomics at workgroup scope order every access (measured 2 % faster than at
// agent scope, which bypasses L1).
__device__ __forceinline__ double ld_acc(const double *p) {
    return __longlong_as_double((long long)__hip_atomic_load(reinterpret_cast<const uint64_t *>(p), __ATOMIC_RELAXED,
                                                             __HIP_MEMORY_SCOPE_WORKGROUP));
}
__device__ __forceinline__ void st_acc(double *p, double v) {
    __hip_atomic_store(reinterpret_cast<uint64_t *>(p), (uint64_t)__double_as_longlong(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <int TPB, int CAP>
__device__ __forceinline__ int64_t large_batch_meta(const Operands &op, const Workspace &ws, int64_t e0, int nbt,
                                                    int64_t *s_bst, double *s_av, int64_t *s_off, int64_t *s_ws64) {
    const int tid = threadIdx.x;
    const int ipt = (nbt + TPB - 1) / TPB;
    const int f = tid * ipt;
    int64_t mylen = 0;
    for (int j = 0; j < ipt; ++j) {
        const int e = f + j;
        if (e < nbt) {
            const int64_t len = ws.elen[e0 + e];
            s_bst[e] = entry_bstart_any(ws, e0 + e);
            s_av[e] = op.aval[e0 + e];
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

// Products s = c0 + tid + j*TPB of the current A-entry batch: B column, B
// value and the batch entry e (its A value is s_av[e]); column -1 past Pb.
// Every lane of the wave calls it (groups of 64 consecutive products; `mark`
// is the wave's 64 ints).
template <int TPB, int CAP, int ITEMS>
__device__ __forceinline__ void load_chunk(const Operands &op, const int64_t *s_bst, const int64_t *s_off, int nbt,
                                           int64_t Pb, int64_t c0, int *mark, int (&col)[ITEMS],
                                           double (&bv)[ITEMS], int (&ent)[ITEMS]) {
    int64_t g[ITEMS];
    int e[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) g[j] = c0 + threadIdx.x + j * TPB;
    group_entries<ITEMS>(s_off, nbt, uni64(c0 + (int64_t)(threadIdx.x - lane_id())), (int64_t)TPB, mark, e);
    // Unconditional loads (products past the batch read B entry 0; the
    // caller only reads columns of products inside it, see chunk_col), so
    // that the compiler waits for exactly one chunk's loads.
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const bool in = g[j] < Pb;
        const int64_t idx = in ? s_bst[e[j]] + (g[j] - s_off[e[j]]) : 0;
        col[j] = op.bcol[idx];
        bv[j] = op.bval[idx];
        ent[j] = in ? e[j] : -1;
    }
}

// Column of a loaded product, -1 past the batch (ent < 0).
__device__ __forceinline__ int chunk_col(int col, int ent) { return ent < 0 ? -1 : col; }

// LDSB: the window's presence bitmap lives in dynamic LDS (lwidth / 8
// bytes), so a head knows whether its column is present before touching HBM
// and loads the running sum of present columns only; else it is the
// workgroup's bitmap in HBM, loaded together with the running sum.
template <int TPB, int CAP, bool LDSB>
__global__ __launch_bounds__(TPB) void k_large_numeric(Operands op, Workspace ws, const int64_t *crowptr,
                                                       int32_t *ccol, double *cval, int64_t skip_le) {
    constexpr int ITEMS = CAP / TPB;
    extern __shared__ uint32_t s_lbits[];
    constexpr int kExtract = 8;      // accumulator loads in flight per thread during extraction
    constexpr int kExtractLds = 16;  // the same with the bitmap in LDS
    __shared__ int64_t s_bst[CAP];
    __shared__ double s_av[CAP];
    __shared__ int64_t s_off[CAP + 1];
    __shared__ uint64_t s_key[2][CAP];  // sorted chunk keys, two chunks in flight
    __shared__ double s_val[2][CAP];
    __shared__ int s_cnt[CAP + 1];
    __shared__ int s_wsi[2 * (TPB / 64) + 2];
    __shared__ int64_t s_ws64[TPB / 64 + 1];
    const int tid = threadIdx.x;
    const RowRec *list = large_list(ws);
    const int count = ws.ctrl->bin_count[kBinLarge];
    const int64_t W = ws.lwidth;
    uint32_t *bits = ws.lbits + (size_t)blockIdx.x * (size_t)(W / 32);
    double *acc = ws.lacc + (size_t)blockIdx.x * (size_t)W;
    unsigned flags = 0;
    unsigned long long stored = 0;
    if (LDSB) {
        for (int64_t q = tid; q < W / 32; q += TPB) s_lbits[q] = 0u;
    }
    __shared__ int s_next;
    __shared__ int s_gmark[TPB / 64][64];  // each wave's entry marks (group_entries)
    for (int i; (i = next_large_row(&ws.ctrl->large_next[1], &s_next)) < count;) {
        const int row = list[i].row;
        if (ws.snnz[row] <= skip_le) continue;  // taken by the compact kernel
        const int64_t a0 = op.aptr[row], a1 = op.aptr[row + 1];
        const int64_t cbase = crowptr[row];
        const int64_t lim = crowptr[row + 1] - cbase;
        int64_t outn = 0;
        for (int64_t w0 = 0; w0 < op.bcols; w0 += W) {
            const int64_t wend = min(w0 + W, (int64_t)op.bcols);
            // This thread's net change of the number of the window's columns
            // whose running sum is exactly zero.
            int zc = 0;
            // The chunk whose heads' running sums are in flight.  It is folded
            // after the next chunk has been bucketed, ranked and placed in the
            // other buffer, so the accumulator round trip overlaps that work.
            int pend_v = 0;  // its valid products (0: none pending)
            int pbuf = 0;    // its sorted keys and values: s_key[pbuf], s_val[pbuf]
            int cbuf = 0;    // buffer of the next chunk
            bool head[ITEMS];
            uint32_t hcol[ITEMS], hword[ITEMS];
            uint64_t hacc[ITEMS];
#pragma unroll
            for (int j = 0; j < ITEMS; ++j) {
                head[j] = false;
                hcol[j] = hword[j] = 0u;
                hacc[j] = 0ull;
            }
            // Each head continues its column's running sum (the accumulator,
            // when an earlier chunk touched it, else its first product) with
            // the column's products in serial order.
            auto fold_pending = [&]() {
#pragma unroll
                for (int j = 0; j < ITEMS; ++j) {
                    if (!head[j]) continue;
                    const int y0 = tid + j * TPB;
                    const uint32_t rel = (uint32_t)((int64_t)(int32_t)hcol[j] - w0);
                    const uint32_t bit = 1u << (rel & 31u);
                    const bool present = (hword[j] & bit) != 0u;
                    const double old = __longlong_as_double((long long)hacc[j]);
                    double v = present ? old + s_val[pbuf][y0] : s_val[pbuf][y0];
                    for (int y = y0 + 1; y < pend_v && (uint32_t)(s_key[pbuf][y] >> 32) == hcol[j]; ++y)
                        v = v + s_val[pbuf][y];
                    st_acc(&acc[rel], v);
                    if (!present) {
                        if (LDSB) atomicOr(&s_lbits[rel >> 5], bit);
                        else atomicOr(&bits[rel >> 5], bit);
                    }
                    zc += (int)(v == 0.0) - (int)(present && old == 0.0);
                    head[j] = false;
                }
                pend_v = 0;
            };
            for (int64_t e0 = a0; e0 < a1; e0 += CAP) {
                const int nbt = (int)min((int64_t)CAP, a1 - e0);
                const int64_t Pb = large_batch_meta<TPB, CAP>(op, ws, e0, nbt, s_bst, s_av, s_off, s_ws64);
                // Each chunk's products are loaded one chunk ahead.
                int pcol[ITEMS], pent[ITEMS];
                double pbv[ITEMS];
                load_chunk<TPB, CAP>(op, s_bst, s_off, nbt, Pb, 0, s_gmark[tid >> 6], pcol, pbv, pent);
                for (int64_t c0 = 0; c0 < Pb; c0 += CAP) {
                    int icol[ITEMS];
                    double ival[ITEMS];
                    bool ok[ITEMS];
                    int cmn = INT_MAX, cmx = INT_MIN, nvalid = 0;
#pragma unroll
                    for (int j = 0; j < ITEMS; ++j) {
                        const int c = chunk_col(pcol[j], pent[j]);
                        ok[j] = c >= w0 && c < wend;  // -1 past the batch
                        icol[j] = ok[j] ? c : 0;
                        ival[j] = ok[j] ? s_av[pent[j]] * pbv[j] : 0.0;
                        if (ok[j]) {
                            cmn = min(cmn, c);
                            cmx = max(cmx, c);
                            ++nvalid;
                        }
                    }
                    load_chunk<TPB, CAP>(op, s_bst, s_off, nbt, Pb, c0 + CAP, s_gmark[tid >> 6], pcol, pbv, pent);
                    const int V = block_sum<TPB>(nvalid, s_wsi);
                    if (V == 0) continue;  // uniform across the block; a pending chunk stays pending
                    uint64_t *const sk = s_key[cbuf];
                    double *const sv = s_val[cbuf];
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
                    block_scan_runs<TPB>(s_cnt, nb, s_wsi);
                    if (tid == 0) s_cnt[nb] = V;
                    __syncthreads();
                    uint64_t key[ITEMS];
#pragma unroll
                    for (int j = 0; j < ITEMS; ++j) {
                        key[j] = ((uint64_t)(uint32_t)icol[j] << 32) | (uint32_t)(tid + j * TPB);
                        if (ok[j]) sk[s_cnt[bk[j]] + pos[j]] = key[j];
                    }
                    __syncthreads();
                    // Rank inside the bucket by (column, serial index): the
                    // item's sorted position.  Keys are distinct.
                    int x[ITEMS];
#pragma unroll
                    for (int j = 0; j < ITEMS; ++j) {
                        x[j] = -1;
                        if (ok[j]) {
                            const int lo = s_cnt[bk[j]], hi = s_cnt[bk[j] + 1];
                            int rank = 0;
                            for (int y = lo; y < hi; ++y) rank += sk[y] < key[j];
                            x[j] = lo + rank;
                        }
                    }
                    __syncthreads();  // every bucket read before the sorted placement
#pragma unroll
                    for (int j = 0; j < ITEMS; ++j) {
                        if (ok[j]) {
                            sk[x[j]] = key[j];
                            sv[x[j]] = ival[j];
                        }
                    }
                    fold_pending();   // the previous chunk's sums have had this chunk's sort to arrive
                    __syncthreads();  // placement done; folded sums and presence bits visible
                    // This chunk's heads (each column's first sorted
                    // position): presence and running-sum loads are issued
                    // now and used when the next chunk has been sorted.
#pragma unroll
                    for (int j = 0; j < ITEMS; ++j) {
                        const int y = tid + j * TPB;
                        head[j] = y < V && (y == 0 || (sk[y - 1] >> 32) != (sk[y] >> 32));
                        hcol[j] = head[j] ? (uint32_t)(sk[y] >> 32) : 0u;
                        const uint32_t rel = head[j] ? (uint32_t)((int64_t)(int32_t)hcol[j] - w0) : 0u;
                        if (LDSB) {
                            hword[j] = head[j] ? s_lbits[rel >> 5] : 0u;
                            const bool present = (hword[j] >> (rel & 31u)) & 1u;
                            hacc[j] = present ? (uint64_t)__double_as_longlong(ld_acc(&acc[rel])) : 0ull;
                        } else {
                            hword[j] = head[j] ? ld_bits(&bits[rel >> 5]) : 0u;
                            hacc[j] = head[j] ? (uint64_t)__double_as_longlong(ld_acc(&acc[rel])) : 0ull;
                        }
                    }
                    pend_v = V;
                    pbuf = cbuf;
                    cbuf ^= 1;
                }
            }
            fold_pending();
            __syncthreads();  // every sum and presence bit of the window in place
            // Extraction of the window in ascending column order, zeros
            // dropped.  When no column of the window sums to exactly zero (or
            // zeros are kept) the entries before a word are the set bits
            // before it, and the counting pass reads the bitmap only.
            const int64_t nwords = (wend - w0 + 31) / 32;
            const int64_t wpt = (nwords + TPB - 1) / TPB;
            const int64_t q0 = min((int64_t)tid * wpt, nwords);
            const int64_t q1 = min(q0 + wpt, nwords);
            const bool all_stored = op.keep_zeros || block_sum<TPB>(zc, s_wsi) == 0;
            int64_t mine = 0;
            if (all_stored) {
                for (int64_t q = q0; q < q1; q += kExtract) {
                    uint32_t wd[kExtract];
#pragma unroll
                    for (int u = 0; u < kExtract; ++u)
                        wd[u] = q + u < q1 ? (LDSB ? s_lbits[q + u] : ld_bits(&bits[q + u])) : 0u;
#pragma unroll
                    for (int u = 0; u < kExtract; ++u) mine += __popc(wd[u]);
                }
            } else {
                for (int64_t q = q0; q < q1; ++q) {
                    uint32_t m = LDSB ? s_lbits[q] : ld_bits(&bits[q]);
                    while (m) {
                        const int b = __ffs((int)m) - 1;
                        m &= m - 1u;
                        if (ld_acc(&acc[q * 32 + b]) != 0.0) ++mine;
                    }
                }
            }
            int64_t wtotal;
            int64_t o = outn + block_excl_scan<TPB>(mine, s_ws64, wtotal);
            if (LDSB) {
                // Set bits are taken across the thread's words in column
                // order, kExtractLds sums loaded per round trip.
                int64_t q = q0;
                uint32_t m = q0 < q1 ? s_lbits[q0] : 0u;
                for (;;) {
                    int rel[kExtractLds];
                    double v[kExtractLds];
#pragma unroll
                    for (int u = 0; u < kExtractLds; ++u) {
                        while (m == 0u && q + 1 < q1) {
                            s_lbits[q] = 0u;
                            m = s_lbits[++q];
                        }
                        rel[u] = -1;
                        if (m) {
                            rel[u] = (int)((q - q0) * 32) + __ffs((int)m) - 1;
                            m &= m - 1u;
                        }
                    }
                    if (rel[0] < 0) break;
#pragma unroll
                    for (int u = 0; u < kExtractLds; ++u)
                        v[u] = rel[u] >= 0 ? ld_acc(&acc[q0 * 32 + rel[u]]) : 0.0;
#pragma unroll
                    for (int u = 0; u < kExtractLds; ++u) {
                        if (rel[u] >= 0 && (op.keep_zeros || v[u] != 0.0)) {
                            if (o < lim) {
                                ccol[cbase + o] = (int32_t)(w0 + q0 * 32 + rel[u]);
                                cval[cbase + o] = v[u];
                            }
                            ++o;
                        }
                    }
                }
                if (q < q1) s_lbits[q] = 0u;
                outn += wtotal;
                __syncthreads();
                continue;
            }
            uint32_t m = q0 < q1 ? ld_bits(&bits[q0]) : 0u;
            for (int64_t q = q0; q < q1; ++q) {
                const uint32_t nxt = q + 1 < q1 ? ld_bits(&bits[q + 1]) : 0u;
                while (m) {
                    int b[kExtract];
                    double v[kExtract];
#pragma unroll
                    for (int u = 0; u < kExtract; ++u) {
                        b[u] = -1;
                        if (m) {
                            b[u] = __ffs((int)m) - 1;
                            m &= m - 1u;
                        }
                    }
#pragma unroll
                    for (int u = 0; u < kExtract; ++u) v[u] = b[u] >= 0 ? ld_acc(&acc[q * 32 + b[u]]) : 0.0;
#pragma unroll
                    for (int u = 0; u < kExtract; ++u) {
                        if (b[u] >= 0 && (op.keep_zeros || v[u] != 0.0)) {
                            if (o < lim) {
                                ccol[cbase + o] = (int32_t)(w0 + q * 32 + b[u]);
                                cval[cbase + o] = v[u];
                            }
                            ++o;
                        }
                    }
                }
                st_bits(&bits[q], 0u);
                m = nxt;
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

// Large rows whose B columns fit one window of at most kCompactCols columns,
// with the presence bitmap the count kernel saved for the row (row i of the
// large-row list at word i * lwidth/32).  A column's rank among the row's
// distinct columns -- the set bits before it, from per-word prefix counts in
// LDS -- indexes a dense accumulator in LDS, so a fold never leaves the CU.
// Each chunk of CAP products (serial order) is sorted by (rank, serial index)
// and each rank's run is added, left to right, to its running sum.  Every
// sum starts at -0.0: -0.0 + p == p for every product p (signed zeros,
// infinities and NaNs included), which is "the first product initialises the
// sum" of the serial order.  Rows with more distinct columns than the
// accumulator holds (acc_cap) are taken in passes over consecutive ranks;
// each pass streams the row's products again.  The sums come out in column
// order, exact zeros dropped unless kept.
template <int TPB, int CAP>
__global__ __launch_bounds__(TPB) void k_large_compact(Operands op, Workspace ws, const uint32_t *lsave, int acc_cap,
                                                       int64_t take_le, const int64_t *crowptr, int32_t *ccol,
                                                       double *cval) {
    constexpr int ITEMS = CAP / TPB;
    extern __shared__ double s_acc[];  // acc_cap sums, then the bitmap and its per-word prefix counts
    __shared__ int64_t s_bst[CAP];
    __shared__ double s_av[CAP];
    __shared__ int64_t s_off[CAP + 1];
    __shared__ uint64_t s_key[CAP];
    __shared__ double s_val[CAP];
    __shared__ int s_cnt[CAP + 1];
    __shared__ int s_wsi[2 * (TPB / 64) + 2];
    __shared__ int64_t s_ws64[TPB / 64 + 1];
    __shared__ int s_next;
    __shared__ int s_gmark[TPB / 64][64];  // each wave's entry marks (group_entries)
    const int tid = threadIdx.x;
    const RowRec *list = large_list(ws);
    const int count = ws.ctrl->bin_count[kBinLarge];
    const int nw = (int)(ws.lwidth / 32);   // words per saved bitmap
    const int nwc = (op.bcols + 31) / 32;   // words covering B's columns (<= nw)
    uint32_t *s_bits = reinterpret_cast<uint32_t *>(s_acc + acc_cap);
    int *s_pre = reinterpret_cast<int *>(s_bits + nw);
    // Each thread owns consecutive bitmap words: prefix counts, extraction.
    const int wpt = (nwc + TPB - 1) / TPB;
    const int q0 = min(tid * wpt, nwc), q1 = min(q0 + wpt, nwc);
    unsigned flags = 0;
    unsigned long long stored = 0;
    for (int i; (i = next_large_row(&ws.ctrl->large_next[2], &s_next)) < count;) {
        const int row = list[i].row;
        if (ws.snnz[row] > take_le) continue;  // taken by the HBM-window kernel
        const int64_t a0 = op.aptr[row], a1 = op.aptr[row + 1];
        const int64_t cbase = crowptr[row];
        const int64_t lim = crowptr[row + 1] - cbase;
        const uint32_t *saved = lsave + (size_t)i * (size_t)nw;
        int mine = 0;
        for (int q = q0; q < q1; ++q) {
            const uint32_t w = saved[q];
            s_bits[q] = w;
            mine += __popc(w);
        }
        int64_t T;
        int64_t run = block_excl_scan<TPB>((int64_t)mine, s_ws64, T);
        for (int q = q0; q < q1; ++q) {
            s_pre[q] = (int)run;
            run += __popc(s_bits[q]);
        }
        __syncthreads();  // bitmap and prefix counts in place
        int64_t outn = 0;
        for (int64_t r0 = 0; r0 < T; r0 += acc_cap) {
            const int racc = (int)min((int64_t)acc_cap, T - r0);
            for (int q = tid; q < racc; q += TPB) s_acc[q] = -0.0;
            __syncthreads();
            for (int64_t e0 = a0; e0 < a1; e0 += CAP) {
                const int nbt = (int)min((int64_t)CAP, a1 - e0);
                const int64_t Pb = large_batch_meta<TPB, CAP>(op, ws, e0, nbt, s_bst, s_av, s_off, s_ws64);
                // Each chunk's products are loaded one chunk ahead.
                int pcol[ITEMS], pent[ITEMS];
                double pbv[ITEMS];
                load_chunk<TPB, CAP>(op, s_bst, s_off, nbt, Pb, 0, s_gmark[tid >> 6], pcol, pbv, pent);
                for (int64_t c0 = 0; c0 < Pb; c0 += CAP) {
                    int rk[ITEMS];
                    double ival[ITEMS];
                    bool ok[ITEMS];
                    int rmn = INT_MAX, rmx = INT_MIN, nvalid = 0;
#pragma unroll
                    for (int j = 0; j < ITEMS; ++j) {
                        const int c = chunk_col(pcol[j], pent[j]);  // -1 past the batch
                        bool in = c >= 0 && c < op.bcols;
                        int r = 0;
                        if (in) {
                            const uint32_t wd = s_bits[c >> 5];
                            if (!((wd >> (c & 31)) & 1u)) flags |= kErrInternal;  // the count saw every column
                            r = s_pre[c >> 5] + __popc(wd & ((1u << (c & 31)) - 1u)) - (int)r0;
                            in = r >= 0 && r < racc;
                        }
                        ok[j] = in;
                        rk[j] = in ? r : 0;
                        ival[j] = in ? s_av[pent[j]] * pbv[j] : 0.0;
                        if (in) {
                            rmn = min(rmn, r);
                            rmx = max(rmx, r);
                            ++nvalid;
                        }
                    }
                    load_chunk<TPB, CAP>(op, s_bst, s_off, nbt, Pb, c0 + CAP, s_gmark[tid >> 6], pcol, pbv, pent);
                    const int V = block_sum<TPB>(nvalid, s_wsi);
                    if (V == 0) continue;  // uniform across the block
                    block_minmax<TPB>(rmn, rmx, s_wsi);
                    int nb = TPB;
                    while (nb < V && nb < CAP) nb <<= 1;
                    const int lognb = 31 - __clz(nb);
                    const uint32_t range = (uint32_t)(rmx - rmn) + 1u;
                    const int lr = range <= 1u ? 0 : 32 - __clz((int)(range - 1u));
                    const int sh = lr > lognb ? lr - lognb : 0;
                    for (int t = tid; t < nb; t += TPB) s_cnt[t] = 0;
                    __syncthreads();
                    int bk[ITEMS], pos[ITEMS];
#pragma unroll
                    for (int j = 0; j < ITEMS; ++j) {
                        bk[j] = 0;
                        pos[j] = 0;
                        if (ok[j]) {
                            bk[j] = (int)((uint32_t)(rk[j] - rmn) >> sh);
                            pos[j] = atomicAdd(&s_cnt[bk[j]], 1);
                        }
                    }
                    __syncthreads();
                    block_scan_runs<TPB>(s_cnt, nb, s_wsi);
                    if (tid == 0) s_cnt[nb] = V;
                    __syncthreads();
                    uint64_t key[ITEMS];
#pragma unroll
                    for (int j = 0; j < ITEMS; ++j) {
                        key[j] = ((uint64_t)(uint32_t)rk[j] << 32) | (uint32_t)(tid + j * TPB);
                        if (ok[j]) s_key[s_cnt[bk[j]] + pos[j]] = key[j];
                    }
                    __syncthreads();
                    // Rank inside the bucket by (rank, serial index): the
                    // item's sorted position.  Keys are distinct.
                    int x[ITEMS];
#pragma unroll
                    for (int j = 0; j < ITEMS; ++j) {
                        x[j] = -1;
                        if (ok[j]) {
                            const int lo = s_cnt[bk[j]], hi = s_cnt[bk[j] + 1];
                            int rank = 0;
                            for (int y = lo; y < hi; ++y) rank += s_key[y] < key[j];
                            x[j] = lo + rank;
                        }
                    }
                    __syncthreads();  // every bucket read before the sorted placement
#pragma unroll
                    for (int j = 0; j < ITEMS; ++j) {
                        if (ok[j]) {
                            s_key[x[j]] = key[j];
                            s_val[x[j]] = ival[j];
                        }
                    }
                    __syncthreads();
                    // Each rank's first sorted position adds the rank's run,
                    // in serial order, to its running sum.
#pragma unroll
                    for (int j = 0; j < ITEMS; ++j) {
                        const int y = tid + j * TPB;
                        if (y < V && (y == 0 || (s_key[y - 1] >> 32) != (s_key[y] >> 32))) {
                            const uint32_t r = (uint32_t)(s_key[y] >> 32);
                            double v = s_acc[r] + s_val[y];
                            for (int z = y + 1; z < V && (uint32_t)(s_key[z] >> 32) == r; ++z) v = v + s_val[z];
                            s_acc[r] = v;
                        }
                    }
                    __syncthreads();  // sums in place; chunk arrays reused next
                }
            }
            // Extraction of ranks [r0, r0 + racc) in column order: the set
            // bits of the thread's words whose rank falls in the pass.
            int cntp = 0;
            for (int q = q0; q < q1; ++q) {
                uint32_t m = s_bits[q];
                int r = s_pre[q] - (int)r0;
                if (r >= racc || r + __popc(m) <= 0) continue;
                for (; m; m &= m - 1u, ++r)
                    if (r >= 0 && r < racc && (op.keep_zeros || s_acc[r] != 0.0)) ++cntp;
            }
            int64_t ptotal;
            int64_t o = outn + block_excl_scan<TPB>((int64_t)cntp, s_ws64, ptotal);
            for (int q = q0; q < q1; ++q) {
                uint32_t m = s_bits[q];
                int r = s_pre[q] - (int)r0;
                if (r >= racc || r + __popc(m) <= 0) continue;
                for (; m; m &= m - 1u, ++r) {
                    if (r < 0 || r >= racc) continue;
                    const double v = s_acc[r];
                    if (op.keep_zeros || v != 0.0) {
                        if (o < lim) {
                            ccol[cbase + o] = q * 32 + (__ffs((int)m) - 1);
                            cval[cbase + o] = v;
                        }
                        ++o;
                    }
                }
            }
            outn += ptotal;
            __syncthreads();  // sums read before the next pass resets them
        }
        if (tid == 0) {
            ws.annz[row] = outn;
            stored += (unsigned long long)outn;
            if (outn > lim || T > lim) flags |= kErrInternal;
        }
    }
    if (tid == 0 && stored) atomicAdd(&ws.ctrl->actual, stored);
    if (flags) atomicOr(&ws.ctrl->flags, (unsigned long long)flags);
}

template <int TPB, int CAP>
void launch_esc(const Operands &op, const Workspace &ws, int bin, int rows, int cus, const int64_t *crowptr,
                int32_t *ccol, double *cval, hipStream_t stream) {
    if (rows <= 0) return;
    auto k = (ws.mode & kWsNarrow) ? k_numeric_esc<TPB, CAP, true> : k_numeric_esc<TPB, CAP, false>;
    const int g = occupancy_grid(k, TPB, cus, rows);
    hipLaunchKernelGGL(k, dim3(g), dim3(TPB), 0, stream, op, ws, bin, crowptr, ccol, cval);
}

// Set once per process by prepare_numeric_kernels.
std::atomic<bool> g_num_wide{false};
std::atomic<int> g_compact_dyn{0};  // dynamic LDS bytes the compact large-row kernel may use

}  // namespace

bool compact_large_available(int64_t lwidth) {
    if (lwidth > kCompactCols) return false;
    // the bitmap and its per-word prefix counts: 2 x lwidth/32 words
    const int64_t acc = ((int64_t)g_compact_dyn.load(std::memory_order_relaxed) - lwidth / 4) / 8;
    return acc >= kCapLargeCompact;
}

// The compact kernel takes rows of at most this many distinct columns: its
// LDS accumulator is then small enough for the partition and fold kernels to
// run beside it on the same CUs (R-MAT x B(65536 columns): 27.7 ms with the
// whole accumulator, 28.7 with no compact rows, 25.5 at 4096, 26.9 at 8192).
#ifndef SPG_COMPACT_TAKE_MAX
#define SPG_COMPACT_TAKE_MAX 4096
#endif
int64_t compact_take_le(int64_t lwidth) {
    if (!compact_large_available(lwidth)) return -1;
    const int dyn = g_compact_dyn.load(std::memory_order_relaxed);
    const int64_t cap = (int64_t)((int)(((int64_t)dyn - lwidth / 4) / 8) & ~63);
    return cap < (int64_t)SPG_COMPACT_TAKE_MAX ? cap : (int64_t)SPG_COMPACT_TAKE_MAX;
}

bool prepare_numeric_kernels(size_t lds_per_block) {
    {
        // The compact large-row kernel's accumulator takes whatever LDS its
        // static arrays leave.
        const void *fk = reinterpret_cast<const void *>(&k_large_compact<kTpbLargeCompact, kCapLargeCompact>);
        hipFuncAttributes attr{};
        int dyn = 0;
        if (hipFuncGetAttributes(&attr, fk) == hipSuccess && lds_per_block > attr.sharedSizeBytes) {
            dyn = (int)(lds_per_block - attr.sharedSizeBytes);
            dyn &= ~255;
            if (hipFuncSetAttribute(fk, hipFuncAttributeMaxDynamicSharedMemorySize, dyn) != hipSuccess) dyn = 0;
        }
        (void)hipGetLastError();
        g_compact_dyn.store(dyn, std::memory_order_relaxed);
    }
    const size_t dyn = (size_t)(kLdsBitsColsWide / 8);
    constexpr int C = kCapLargeWideNum;
    // static arrays: batch metadata 3 x 8 B, two chunk buffers 2 x 16 B, bucket counts 4 B per product
    const size_t stat = (size_t)C * (3 * 8 + 2 * 16 + 4) + 64 * 8 + 64;
    if (lds_per_block < dyn + stat) return false;
    const hipError_t e =
        hipFuncSetAttribute(reinterpret_cast<const void *>(&k_large_numeric<kTpbLargeWideNum, kCapLargeWideNum, true>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    g_num_wide.store(true, std::memory_order_relaxed);
    return true;
}

void launch_numeric(const Operands &op, const Workspace &ws, const int *bin_counts, unsigned bins, int cus,
                    const int64_t *crowptr, int32_t *ccol, double *cval, hipStream_t stream, const uint32_t *lsave,
                    hipStream_t side, hipEvent_t fork, hipEvent_t join, const PartState *part) {
    if (bins & (1u << kBinW1))
        launch_esc<64, kCapW1>(op, ws, kBinW1, bin_counts[kBinW1], cus, crowptr, ccol, cval, stream);
    if (bins & (1u << kBinW2))
        launch_esc<64, kCapW2>(op, ws, kBinW2, bin_counts[kBinW2], cus, crowptr, ccol, cval, stream);
    if (bins & (1u << kBinW3))
        launch_esc<64, kCapW3>(op, ws, kBinW3, bin_counts[kBinW3], cus, crowptr, ccol, cval, stream);
    if (bins & (1u << kBinW4))
        launch_esc<kTpbW4, kCapW4>(op, ws, kBinW4, bin_counts[kBinW4], cus, crowptr, ccol, cval, stream);
    if (bins & (1u << kBinBlock))
        launch_esc<kTpbBlock, kCapBlock>(op, ws, kBinBlock, bin_counts[kBinBlock], cus, crowptr, ccol, cval, stream);
    if ((bins & (1u << kBinLarge)) && bin_counts[kBinLarge] > 0) {
        // Rows whose distinct columns fit the compact kernel's accumulator
        // (bitmaps saved by the count) fold in LDS; the others are
        // partitioned by column window (`part`) or fold onto the HBM
        // windows.  Each kernel skips the other's rows.
        const bool compact = lsave && compact_large_available(ws.lwidth);
        const int acc_cap = compact ? (int)compact_take_le(ws.lwidth) : 0;
        const int64_t skip_le = compact ? acc_cap : -1;
        const bool forked = compact && side && fork && join;
        if (forked) {
            (void)hipEventRecord(fork, stream);
            (void)hipStreamWaitEvent(side, fork, 0);
        }
        const int g = bin_counts[kBinLarge] < ws.lblocks ? bin_counts[kBinLarge] : ws.lblocks;
        if (part)
            launch_part(op, ws, *part, bin_counts[kBinLarge], cus, crowptr, ccol, cval, stream);
        else if (ws.lwidth <= kLdsBitsCols)
            hipLaunchKernelGGL((k_large_numeric<kTpbLarge, kCapLarge, true>), dim3(g), dim3(kTpbLarge),
                               (unsigned)(ws.lwidth / 8), stream, op, ws, crowptr, ccol, cval, skip_le);
        else if (ws.lwidth <= kLdsBitsColsWide && g_num_wide.load(std::memory_order_relaxed))
            hipLaunchKernelGGL((k_large_numeric<kTpbLargeWideNum, kCapLargeWideNum, true>), dim3(g),
                               dim3(kTpbLargeWideNum), (unsigned)(ws.lwidth / 8), stream, op, ws, crowptr, ccol, cval,
                               skip_le);
        else
            hipLaunchKernelGGL((k_large_numeric<kTpbLargeHbm, kCapLargeHbm, false>), dim3(g), dim3(kTpbLargeHbm), 0,
                               stream, op, ws, crowptr, ccol, cval, skip_le);
        if (compact) {
            // One workgroup per CU (its accumulator fills the CU's LDS).
            const int gc = bin_counts[kBinLarge] < cus ? bin_counts[kBinLarge] : cus;
            hipLaunchKernelGGL((k_large_compact<kTpbLargeCompact, kCapLargeCompact>), dim3(gc),
                               dim3(kTpbLargeCompact), (unsigned)((size_t)acc_cap * 8 + (size_t)(ws.lwidth / 4)),
                               forked ? side : stream, op, ws, lsave, acc_cap, (int64_t)acc_cap, crowptr, ccol, cval);
            if (forked) {
                (void)hipEventRecord(join, side);
                (void)hipStreamWaitEvent(stream, join, 0);
            }
        }
    }
}

}  // namespace spg
