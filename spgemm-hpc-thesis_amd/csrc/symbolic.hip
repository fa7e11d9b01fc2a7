// symbolic.hip -- exact structural count (distinct output columns) per row.
//
// The count does not depend on summation order, so small and medium rows use
// an open-addressing hash set of column indices in LDS (table two or four times
// the row's product bound, so probing always ends).  A entries are expanded
// "entry-major": with L the power of two >= the longest referenced B row, each
// group of L lanes streams one B row, so there is no per-product search, reads
// of a B row are contiguous, and all of a row's column loads are in flight
// before the first insertion.  Insertions then run in lockstep probe rounds
// with every register slot's compare-and-swap issued back to back.
// Rows whose columns span fewer than 32 x the table's words (at least 65536
// columns for rows of up to 512 products) count with a bitmap over their
// column range in the table's storage instead: one LDS OR per product, no
// probing.  Large rows set bits in a per-workgroup column bitmap (in LDS for
// windows of up to 2^18 columns, else in HBM), window by window, and count
// the bits they newly set.
#include "device_common.hpp"

#include <atomic>

namespace spg {
namespace {

using namespace dev;

constexpr int32_t kEmptyKey = INT_MAX;  // never a valid column (col < cols <= INT_MAX)
constexpr int kLargeUnroll = 4;         // products in flight per thread on the large-row path

__device__ __forceinline__ uint32_t hash_col(int32_t c) { return (uint32_t)c * 2654435761u; }

template <int TPB, int CAP>
struct HashLds {
    static constexpr int FILL = CAP <= 256 ? 4 : 2;  // table size / product bound
    static constexpr int TMAX = FILL * CAP;
    // Words of the bitmap path: at least 2048 (columns spanning < 65536) from
    // 512-product rows on, so rows of products with up to 64 K columns never probe.
    static constexpr int BMW = CAP >= 512 && TMAX < 2048 ? 2048 : TMAX;
    static constexpr int TABW = (BMW > TMAX + TPB ? BMW : TMAX + TPB);
    // A row's metadata is dead once its column loads are issued, and the
    // table (or bitmap) is used only after that: they share storage, so more
    // rows stay resident per CU (bin W3: 8 instead of 14.6 KB per workgroup).
    union {
        struct {
            int64_t bst[CAP];
            int off[CAP + 1];
        };
        int32_t tab[TABW];  // table (dummy word per thread at TMAX + tid), or the bitmap
    };
    int wsi[2 * (TPB / 64) + 2];
};

// One row of the software pipeline: the row's metadata comes from registers
// loaded one step earlier (pcur), and the next row's record and metadata are
// issued into the other register set (pnext) before this row's gathers.
template <int TPB, int CAP, bool NARROW>
__device__ __forceinline__ void hash_step(const Operands &op, const Workspace &ws, const RowRec *list, int count,
                                          int G, int &i, RowRec &cur, const MetaPre &pcur, MetaPre &pnext,
                                          int &wnext, HashLds<TPB, CAP> &sh, unsigned long long &sym,
                                          unsigned &flags) {
    constexpr int ITEMS = CAP / TPB;
    constexpr int FILL = HashLds<TPB, CAP>::FILL;
    constexpr int TMAX = HashLds<TPB, CAP>::TMAX;
    constexpr int BMW = HashLds<TPB, CAP>::BMW;
    const int tid = threadIdx.x;
    int P, L;
    meta_to_lds<TPB, false, NARROW>(op, ws, cur, pcur, cur.na <= TPB, sh.bst, nullptr, sh.off, sh.wsi, P, L);
    const RowRec nxt = i + G < count ? rec_decode(wnext) : cur;
    wnext = rec_word(list, min(i + 2 * G, count - 1));
    pnext = meta_prefetch<false, NARROW>(op, ws, nxt);
    if (P > CAP) flags |= kErrInternal;  // cannot happen: binning used the same lengths
    if (P <= CAP) {
        if (tid == 0) sh.off[cur.na] = P;
        __syncthreads();
        // All of the row's column loads are issued before any insertion.
        int e[ITEMS], t[ITEMS], icol[ITEMS];
        bool pend[ITEMS];
        slot_products<TPB, ITEMS>(cur.na, P, L, sh.off, e, t, pend);
        int64_t idx[ITEMS];
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) idx[j] = pend[j] ? sh.bst[e[j]] + t[j] : 0;
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) icol[j] = op.bcol[idx[j]];
        int mn = INT_MAX, mx = INT_MIN;
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            mn = pend[j] ? min(mn, icol[j]) : mn;
            mx = pend[j] ? max(mx, icol[j]) : mx;
        }
        block_minmax<TPB>(mn, mx, sh.wsi);
        int fresh = 0;
        if ((int64_t)mx - (int64_t)mn < (int64_t)BMW * 32) {
            // The row's columns fit a bitmap over [mn, mx] in the table's
            // storage: one OR per product, and a product counts when it sets
            // its bit.  Touched words are zeroed first (the storage held the
            // row's metadata, read by every thread before the barriers of
            // the min/max above).
            uint32_t wd[ITEMS], bit[ITEMS];
#pragma unroll
            for (int j = 0; j < ITEMS; ++j) {
                const uint32_t rel = pend[j] ? (uint32_t)(icol[j] - mn) : 0u;
                wd[j] = rel >> 5;
                bit[j] = 1u << (rel & 31u);
            }
#pragma unroll
            for (int j = 0; j < ITEMS; ++j)
                if (pend[j]) sh.tab[wd[j]] = 0;
            __syncthreads();
#pragma unroll
            for (int j = 0; j < ITEMS; ++j) {
                if (pend[j]) {
                    const uint32_t old = atomicOr(reinterpret_cast<uint32_t *>(&sh.tab[wd[j]]), bit[j]);
                    fresh += (old & bit[j]) ? 0 : 1;
                }
            }
            const int total = block_sum<TPB>(fresh, sh.wsi);
            if (tid == 0) ws.snnz[cur.row] = total;
            sym += total;
            __syncthreads();
            cur = nxt;
            i += G;
            return;
        }
        int T = 64;
        while (T < FILL * P && T < TMAX) T <<= 1;
        T = uni(T);
        const int logT = 31 - __clz(T);
        uint32_t h[ITEMS];
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) h[j] = hash_col(icol[j]) >> (32 - logT);
        // The table's T words start empty (the storage held the row's
        // metadata, read by every thread before the barriers above).
        for (int q = tid; q < T; q += TPB) sh.tab[q] = kEmptyKey;
        __syncthreads();
        // Linear probing in lockstep rounds, branch-free: a settled slot aims
        // its compare-and-swap at this thread's private dummy word with a
        // compare value that never matches.  T >= 2P > distinct columns, so
        // every item settles within T rounds.
        for (int round = 0; round < T; ++round) {
            int32_t old[ITEMS];
#pragma unroll
            for (int j = 0; j < ITEMS; ++j)
                old[j] = atomicCAS(pend[j] ? &sh.tab[h[j]] : &sh.tab[TMAX + tid], pend[j] ? kEmptyKey : INT_MIN,
                                   icol[j]);
            int left = 0;
#pragma unroll
            for (int j = 0; j < ITEMS; ++j) {
                const bool ins = pend[j] & (old[j] == kEmptyKey);
                const bool dup = pend[j] & (old[j] == icol[j]);
                fresh += ins;
                pend[j] = pend[j] & !ins & !dup;
                h[j] = pend[j] ? ((h[j] + 1u) & (uint32_t)(T - 1)) : h[j];
                left += pend[j];
            }
            if (!__syncthreads_or(left)) break;
        }
        const int total = block_sum<TPB>(fresh, sh.wsi);
        if (tid == 0) ws.snnz[cur.row] = total;
        sym += total;
    }
    __syncthreads();
    cur = nxt;
    i += G;
}

// Bins W3 and W4 keep six waves per SIMD (80 VGPRs, no spills; W4's LDS
// allows seven 256-thread workgroups per CU, so registers bound its resident rows).
template <int TPB, int CAP>
constexpr int hash_min_waves() { return CAP == kCapW3 || CAP == kCapW4 ? 6 : 1; }

template <int TPB, int CAP, bool NARROW>
__global__ __launch_bounds__(TPB, (hash_min_waves<TPB, CAP>())) void k_symbolic_hash(Operands op, Workspace ws,
                                                                                     int binid) {
    __shared__ HashLds<TPB, CAP> sh;
    const int tid = threadIdx.x;
    const RowRec *list = bin_list(ws, binid);
    const int count = ws.ctrl->bin_count[binid];
    const int G = (int)gridDim.x;
    unsigned flags = 0;
    unsigned long long sym = 0;
    int i = (int)blockIdx.x;
    if (i < count) {
        // Two-step unrolled loop with ping-pong prefetch registers (pa, pb), so
        // no register that an outstanding load writes is ever copied.
        RowRec cur = rec_decode(rec_word(list, i));
        MetaPre pa = meta_prefetch<false, NARROW>(op, ws, cur), pb;
        int wnext = rec_word(list, min(i + G, count - 1));
        for (;;) {
            if (i >= count) break;
            hash_step<TPB, CAP, NARROW>(op, ws, list, count, G, i, cur, pa, pb, wnext, sh, sym, flags);
            if (i >= count) break;
            hash_step<TPB, CAP, NARROW>(op, ws, list, count, G, i, cur, pb, pa, wnext, sh, sym, flags);
        }
    }
    if (tid == 0 && sym) atomicAdd(&ws.ctrl->sym_total, sym);
    if (flags) atomicOr(&ws.ctrl->flags, (unsigned long long)flags);
}

// A batch of nbt A entries: each entry's B row start and exclusive product
// offset (s_off[nbt] = the batch's products, returned), and the longest B
// row of the batch (maxlen).
template <int TPB, int CAP>
__device__ __forceinline__ int64_t large_batch_meta(const Operands &op, const Workspace &ws, int64_t e0, int nbt,
                                                    int64_t *s_bst, int64_t *s_off, int64_t *s_ws64, int *s_wsi,
                                                    int &maxlen) {
    const int tid = threadIdx.x;
    const int ipt = (nbt + TPB - 1) / TPB;
    const int f = tid * ipt;
    int64_t mylen = 0;
    int mx = 0;
    for (int j = 0; j < ipt; ++j) {
        const int e = f + j;
        if (e < nbt) {
            const int len = ws.elen[e0 + e];
            s_bst[e] = entry_bstart_any(ws, e0 + e);
            s_off[e] = len;
            mylen += len;
            mx = max(mx, len);
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
    if (tid == 0) s_off[nbt] = Pb;
    maxlen = block_max<TPB>(mx, s_wsi);  // its barriers also publish s_off
    return Pb;
}

// LDSB: the window's bitmap lives in dynamic LDS (lwidth / 8 bytes) and the
// claims are LDS atomics; else it is the workgroup's bitmap in HBM.  With
// `lsave` (LDSB, B's columns in one window) each row's bitmap is also saved,
// row i of the large-row list at word i * lwidth/32, for the compact value
// kernel.  With `pw` (LDSB, B's columns in one window; partitioned large
// rows) each row's products and distinct columns per window of kPartCols
// columns go to pw[i * nwin + k] and pd[i * nwin + k], and the rows with
// snnz > part_skip_le add their product counts into the control block.
//
// Rows of at least heavy_ge products among the first heavy_n of the list
// (which is sorted heaviest first) can be counted by a separate launch of
// heavy_n workgroups, one per row (heavy_only: workgroup b takes list row b
// if it is that heavy), so that each of them has the waves of a 1024-thread
// workgroup; the launch over the other rows (heavy_only = 0) skips them.
template <int TPB, int CAP, bool LDSB>
__global__ __launch_bounds__(TPB) void k_large_symbolic(Operands op, Workspace ws, uint32_t *lsave, PartState ps,
                                                        int64_t heavy_ge, int heavy_n, int heavy_only) {
    constexpr int UNROLL = LDSB ? 2 * kLargeUnroll : kLargeUnroll;  // 4x measured slower (registers)
    extern __shared__ uint32_t s_lbits[];
    __shared__ int64_t s_bst[CAP];
    __shared__ int64_t s_off[CAP + 1];
    __shared__ int64_t s_ws64[TPB / 64 + 1];
    __shared__ uint32_t s_pw[kPartMaxWin];  // products per partition window (32-bit: rows past 2^31 are not partitioned)
    __shared__ int s_pd[kPartMaxWin];       // distinct columns per partition window (zero between rows)
    // each wave's entry marks (group_entries), all groups of a step at once
    // where the LDS allows it
    constexpr int MARKS = TPB >= 1024 ? 2 : UNROLL;
    __shared__ int s_gmark[TPB / 64][MARKS * 64];
    __shared__ int s_wsi[TPB / 64];
    const int tid = threadIdx.x, lane = tid & 63;
    const RowRec *list = large_list(ws);
    const int count = ws.ctrl->bin_count[kBinLarge];
    if (count == 0) return;  // launched before the host knows the bins; no row takes a ticket
    const int64_t W = ws.lwidth;
    uint32_t *bits = LDSB ? nullptr : ws.lbits + (size_t)blockIdx.x * (size_t)(W / 32);  // HBM bitmap: grid is lblocks
    unsigned flags = 0;
    unsigned long long sym = 0;
    unsigned long long part_sum = 0, part_max = 0;
    const bool part = LDSB && ps.pw != nullptr;
    const int nwin = ps.nwin;
    static_assert(!LDSB || CAP == kCapLarge, "partition batches are the count's batches");
    if (LDSB) {
        for (int64_t q = tid; q < W / 32; q += TPB) s_lbits[q] = 0u;
    }
    if (part) {
        for (int k = tid; k < kPartMaxWin; k += TPB) {
            s_pw[k] = 0u;
            s_pd[k] = 0;
        }
    }
    __shared__ int s_next;
    // The next row's ticket is taken when a row starts (thread 0), so that the
    // counter's round trip overlaps the row's work.
    int ticket = 0;
    if (!heavy_only && tid == 0) ticket = atomicAdd(&ws.ctrl->large_next[0], 1);
    for (int i0 = heavy_only ? (int)blockIdx.x : 0;;) {
        int i;
        if (heavy_only) {
            if (i0 < 0 || i0 >= count || i0 >= heavy_n || ws.ub[list[i0].row] < heavy_ge) break;
            i = i0;
            i0 = -1;  // one row per workgroup
        } else {
            if (tid == 0) s_next = ticket;
            __syncthreads();
            i = s_next;
            __syncthreads();  // every thread has read the slot before it is reused
            if (i >= count) break;
            if (tid == 0) ticket = atomicAdd(&ws.ctrl->large_next[0], 1);
            if (i < heavy_n && ws.ub[list[i].row] >= heavy_ge) continue;  // the heavy launch's
        }
        const int row = list[i].row;
        const int64_t a0 = op.aptr[row], a1 = op.aptr[row + 1];
        int64_t cnt = 0;
        for (int64_t w0 = 0; w0 < op.bcols; w0 += W) {
            const int64_t wend = min(w0 + W, (int64_t)op.bcols);
            for (int64_t e0 = a0; e0 < a1; e0 += CAP) {
                const int nbt = (int)min((int64_t)CAP, a1 - e0);
                if (part) {  // the window counts before this batch (every claim of the previous one is done)
                    const int64_t u = ps.segbase[count + 1 + i] + (e0 - a0) / CAP;
                    for (int k = tid; k < nwin; k += TPB) ps.pseg[(size_t)u * (size_t)nwin + k] = s_pw[k];
                }
                int maxlen;
                const int64_t Pb = large_batch_meta<TPB, CAP>(op, ws, e0, nbt, s_bst, s_off, s_ws64, s_wsi, maxlen);
                // Batches whose B rows all have at most 64 entries are taken
                // entry-major (EM): Lp lanes (a power of two >= the longest
                // row) per A entry, G = 64 / Lp entries per wave and item, no
                // entry search.  Item u of step k of wave w holds entries
                // ((k * UNROLL + u) * NW + w) * G + lane / Lp.
                const bool em = LDSB && maxlen <= 64;
                int logLp = 0;
                while ((1 << logLp) < maxlen) ++logLp;
                const int G = 64 >> logLp;
                constexpr int64_t kStep = (int64_t)TPB * UNROLL;
                // UNROLL products per thread per step: every column load,
                // then every bit claim, in flight together.  With the bitmap
                // in LDS the next step's column loads are issued before this
                // step's claims (LDS atomics do not wait on them).
                // Lane l of wave w holds products s0 + u * TPB (u < UNROLL),
                // s0 = 64 w + l + k * TPB * UNROLL: groups of 64 consecutive
                // products, whose entries group_entries finds together.
                // The loads are unconditional (products past the batch read
                // B entry 0, the batch has products): with no branch around
                // them the compiler waits for exactly the previous step's
                // loads, not for the next step's too.
                const int wv = tid >> 6;
                auto load_step = [&](int64_t k, int (&c)[UNROLL], bool (&v)[UNROLL]) {
                    if (em) {
#pragma unroll
                        for (int u = 0; u < UNROLL; ++u) {
                            const int64_t e = ((k * UNROLL + u) * (TPB / 64) + wv) * G + (lane >> logLp);
                            const int sub = lane & ((1 << logLp) - 1);
                            const int ec = (int)min(e, (int64_t)nbt - 1);
                            v[u] = e < nbt && sub < (int)(s_off[ec + 1] - s_off[ec]);
                            c[u] = op.bcol[v[u] ? s_bst[ec] + sub : 0];
                        }
                    } else {
                        const int64_t s0 = k * kStep + tid;
                        int64_t sx[UNROLL];
                        int e[UNROLL];
#pragma unroll
                        for (int u = 0; u < UNROLL; ++u) sx[u] = s0 + (int64_t)u * TPB;
                        group_entries<UNROLL, MARKS>(s_off, nbt, uni64(s0 - lane), (int64_t)TPB, s_gmark[wv], e);
#pragma unroll
                        for (int u = 0; u < UNROLL; ++u) {
                            v[u] = sx[u] < Pb;
                            c[u] = op.bcol[v[u] ? s_bst[e[u]] + (sx[u] - s_off[e[u]]) : 0];
                        }
                    }
                };
                // steps of the wave: while its first product (EM: first entry)
                // is inside the batch (uniform per wave)
                const int64_t nsteps =
                    em ? ((int64_t)nbt + (int64_t)G * (TPB / 64) * UNROLL - 1) / ((int64_t)G * (TPB / 64) * UNROLL)
                       : (Pb + kStep - 1) / kStep;
                int cn[UNROLL];
                bool vn[UNROLL];
                if (LDSB) load_step(0, cn, vn);
                for (int64_t k = 0; k < nsteps; ++k) {
                    int c[UNROLL];
                    bool cv[UNROLL];
                    if (LDSB) {
#pragma unroll
                        for (int u = 0; u < UNROLL; ++u) {
                            c[u] = cn[u];
                            cv[u] = vn[u];
                        }
                        if (k + 1 < nsteps) load_step(k + 1, cn, vn);
                    } else {
                        load_step(k, c, cv);
                    }
                    uint32_t old[UNROLL], bit[UNROLL];
                    bool in[UNROLL];
#pragma unroll
                    for (int u = 0; u < UNROLL; ++u) {
                        in[u] = cv[u] && c[u] >= w0 && c[u] < wend;
                        const uint32_t rel = in[u] ? (uint32_t)(c[u] - w0) : 0u;
                        bit[u] = 1u << (rel & 31u);
                        if (LDSB) {
                            // no return value to wait for: the row's count is
                            // the bitmap's popcount when it is cleared
                            if (in[u]) atomicOr(&s_lbits[rel >> 5], bit[u]);
                        } else {
                            old[u] = in[u] ? atomicOr(&bits[rel >> 5], bit[u]) : bit[u];
                        }
                    }
                    if (!LDSB) {
#pragma unroll
                        for (int u = 0; u < UNROLL; ++u) cnt += in[u] && !(old[u] & bit[u]);
                    }
                    if (part) {
                        // consecutive lanes hold consecutive products (a B
                        // row's columns ascend): one add per run of lanes in
                        // the same window (lanes past the batch's last
                        // product end runs)
                        const uint64_t active = __ballot(true);
#pragma unroll
                        for (int u = 0; u < UNROLL; ++u) {
                            const uint32_t wk = in[u] ? (uint32_t)(c[u] - w0) >> kPartLog : 0xffffffffu;
                            const uint32_t prev = (uint32_t)dpp_i<0x138, 0xf>(0, (int)wk);  // lane - 1 (wave_shr 1)
                            const bool start = in[u] && (lane == 0 || prev != wk);
                            const uint64_t ends = __ballot(start) | __ballot(!in[u]) | ~active;
                            if (start) {
                                const uint64_t rest = lane == 63 ? 0ull : ends >> (lane + 1);
                                const uint32_t len = rest ? (uint32_t)__builtin_ctzll(rest) + 1u : (uint32_t)(64 - lane);
                                atomicAdd(&s_pw[wk], len);
                            }
                        }
                    }
                }
                __syncthreads();  // batch metadata is overwritten next
            }
            const int64_t nwords = (wend - w0 + 31) / 32;
            uint32_t *save = LDSB && lsave ? lsave + (size_t)i * (size_t)(W / 32) : nullptr;
            if (LDSB) {
                // The row's count is the bitmap's popcount, taken as it is
                // cleared; with partitions each wave's 64 words lie in one
                // window of kPartCols / 32 words, whose distinct columns get
                // the wave's sum.
                static_assert(kPartCols % (32 * 64) == 0, "a wave's 64 bitmap words lie in one partition window");
                for (int64_t q0 = tid - lane; q0 < nwords; q0 += TPB) {
                    const int64_t q = q0 + lane;
                    const uint32_t wd = q < nwords ? s_lbits[q] : 0u;
                    const int pc = __popc(wd);
                    cnt += pc;
                    if (q < nwords) {
                        if (save) save[q] = wd;
                        s_lbits[q] = 0u;
                    }
                    if (part) {
                        const int d = wave_last(wave_incl_sum_dpp(pc));
                        if (lane == 0 && d) atomicAdd(&s_pd[q0 / (kPartCols / 32)], d);
                    }
                }
                if (part) {
                    __syncthreads();  // every window's sum is in
                    for (int k = tid; k < nwin; k += TPB) {
                        ps.pw[(size_t)i * (size_t)nwin + k] = (int64_t)s_pw[k];
                        ps.pd[(size_t)i * (size_t)nwin + k] = s_pd[k];
                        s_pw[k] = 0u;
                        s_pd[k] = 0;
                    }
                }
                // The bitmap is LDS: the next window's or row's claims are
                // ordered by the barrier alone (a device-scope fence would
                // write back and invalidate the L2 at every row).
                __syncthreads();
            } else {
                for (int64_t q = tid; q < nwords; q += TPB) st_bits(&bits[q], 0u);
                __threadfence();  // the HBM bitmap's clears reach the L2 before the next claims
                __syncthreads();
            }
        }
        const int64_t total = block_sum<TPB>(cnt, s_ws64);
        if (tid == 0) ws.snnz[row] = total;
        sym += (unsigned long long)total;
        if (part && total > ps.skip_le) {
            const unsigned long long p = (unsigned long long)ws.ub[row];
            part_sum += p;
            part_max = p > part_max ? p : part_max;
        }
    }
    if (tid == 0 && sym) atomicAdd(&ws.ctrl->sym_total, sym);
    if (tid == 0 && part_sum) {
        atomicAdd(&ws.ctrl->part_products, part_sum);
        atomicMax(&ws.ctrl->part_max_row, part_max);
    }
    if (flags) atomicOr(&ws.ctrl->flags, (unsigned long long)flags);
}

template <int TPB, int CAP>
void launch_hash(const Operands &op, const Workspace &ws, int bin, int cus, hipStream_t stream) {
    auto k = (ws.mode & kWsNarrow) ? k_symbolic_hash<TPB, CAP, true> : k_symbolic_hash<TPB, CAP, false>;
    const int g = occupancy_grid(k, TPB, cus, op.arows);
    hipLaunchKernelGGL(k, dim3(g), dim3(TPB), 0, stream, op, ws, bin);
}

// Set once per process by prepare_kernels: the device gives a workgroup
// enough LDS for the wide-window count.
std::atomic<bool> g_wide_lds{false};

}  // namespace

hipError_t prepare_kernels(size_t lds_per_block) {
    // The wide-window count keeps a bitmap of up to 2^20 columns (128 KB) in
    // dynamic LDS beside its static arrays; without that much LDS per
    // workgroup it is not used.
    const size_t dyn = (size_t)(kLdsBitsColsWide / 8);
    hipFuncAttributes attr{};
    const bool have =
        hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(&k_large_symbolic<kTpbLargeWide, kCapLarge, true>)) ==
        hipSuccess;
    if (!have) (void)hipGetLastError();
    const size_t stat = have ? (size_t)attr.sharedSizeBytes : lds_per_block;  // unknown: not used
    bool ok = lds_per_block >= dyn + stat;
    if (ok) {
        const hipError_t e = hipFuncSetAttribute(
            reinterpret_cast<const void *>(&k_large_symbolic<kTpbLargeWide, kCapLarge, true>),
            hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            ok = false;
        }
    }
    g_wide_lds.store(ok, std::memory_order_relaxed);
    (void)prepare_numeric_kernels(lds_per_block);
    return hipSuccess;
}

bool part_available(int64_t bcols, int64_t lwidth) {
    if (bcols <= 0 || bcols > lwidth || part_windows(bcols) > kPartMaxWin) return false;
    return lwidth <= kLdsBitsCols || (lwidth <= kLdsBitsColsWide && g_wide_lds.load(std::memory_order_relaxed));
}

void launch_symbolic(const Operands &op, const Workspace &ws, unsigned bins, int cus, hipStream_t stream,
                     uint32_t *lsave, const PartState *part, int nlarge, hipStream_t side, hipEvent_t fork,
                     hipEvent_t join) {
    if (op.arows <= 0) return;
    if (bins & (1u << kBinW1)) launch_hash<64, kCapW1>(op, ws, kBinW1, cus, stream);
    if (bins & (1u << kBinW2)) launch_hash<64, kCapW2>(op, ws, kBinW2, cus, stream);
    if (bins & (1u << kBinW3)) launch_hash<kSymTpbW3, kCapW3>(op, ws, kBinW3, cus, stream);
    if (bins & (1u << kBinW4)) launch_hash<kSymTpbW4, kCapW4>(op, ws, kBinW4, cus, stream);
    if (bins & (1u << kBinBlock)) launch_hash<kSymTpbBlock, kCapBlock>(op, ws, kBinBlock, cus, stream);
    if ((bins & (1u << kBinLarge)) && ws.lblocks > 0) {
        PartState ps{};  // no partition counts unless asked for and possible
        ps.skip_le = -1;
        if (part && part->pw && part->pd && part->pseg && part->segbase && part_available(op.bcols, ws.lwidth))
            ps = *part;
        if (ws.lwidth <= kLdsBitsCols) {
            // The bitmap is in LDS, so the grid is not bound to the per-block
            // HBM windows: every workgroup that fits takes rows.
            const size_t dyn = (size_t)(ws.lwidth / 8);
            uint32_t *save = op.bcols <= ws.lwidth ? lsave : nullptr;
            // With the number of large rows known, the heaviest (a row of
            // 1.5 M products otherwise holds one 256-thread workgroup for most
            // of the kernel) are counted by 1024-thread workgroups of their
            // own, one per row.  They are few, so with a side stream they run
            // beside the other rows' count (which skips them) instead of
            // before it.
            int heavy_n = 0;
            const bool forked = nlarge > 0 && side && fork && join;
            if (nlarge > 0 && g_wide_lds.load(std::memory_order_relaxed)) {
                heavy_n = nlarge < kCountHeavyRows ? nlarge : kCountHeavyRows;
                if (forked) {
                    (void)hipEventRecord(fork, stream);
                    (void)hipStreamWaitEvent(side, fork, 0);
                }
                hipLaunchKernelGGL((k_large_symbolic<kTpbLargeWide, kCapLarge, true>), dim3(heavy_n),
                                   dim3(kTpbLargeWide), (unsigned)dyn, forked ? side : stream, op, ws, save, ps,
                                   kCountHeavy, heavy_n, 1);
                if (forked) (void)hipEventRecord(join, side);
            }
            const int g = occupancy_grid(k_large_symbolic<kTpbLarge, kCapLarge, true>, kTpbLarge, cus, 0, dyn);
            hipLaunchKernelGGL((k_large_symbolic<kTpbLarge, kCapLarge, true>), dim3(g), dim3(kTpbLarge),
                               (unsigned)dyn, stream, op, ws, save, ps, kCountHeavy, heavy_n, 0);
            if (forked && heavy_n > 0) (void)hipStreamWaitEvent(stream, join, 0);
        }
        else if (ws.lwidth <= kLdsBitsColsWide && g_wide_lds.load(std::memory_order_relaxed))
            hipLaunchKernelGGL((k_large_symbolic<kTpbLargeWide, kCapLarge, true>), dim3(ws.lblocks),
                               dim3(kTpbLargeWide), (unsigned)(ws.lwidth / 8), stream, op, ws,
                               (uint32_t *)nullptr, ps, (int64_t)0, 0, 0);
        else
            hipLaunchKernelGGL((k_large_symbolic<kTpbLargeHbm, kCapLargeHbm, false>), dim3(ws.lblocks),
                               dim3(kTpbLargeHbm), 0, stream, op, ws, (uint32_t *)nullptr, PartState{}, (int64_t)0, 0,
                               0);
    }
}

}  // namespace spg
