// large_part.hip -- large rows by partition: products moved to HBM scratch by
// column window, then every window folded in an LDS accumulator.
//
// Semantics are the reference's (src/baselines/spgemm_exchange.cpp:274-309):
// per output row, products a*b in serial order (A entries in stored order,
// then B entries in stored order), a column's first product starts its sum and
// later ones are added left to right, columns ascending, sums equal to 0.0
// dropped.
//
// A large row's columns are cut into windows of kPartCols columns.  The count
// kernel has already counted each window's products and distinct columns.
//   1. k_large_part takes units of one batch of kCapLarge A entries of a row
//      (rows heaviest first; the count kernel recorded every window's products
//      before each batch, so a batch knows where its products go).  It streams
//      the batch's products in serial order and moves each one -- its column
//      inside the window (16 bits) and its product value -- to the window's
//      region of the scratch.  The split is stable: each wave takes a
//      contiguous range of the serial order, and a lane's position among the
//      wave's items of the same window comes from ballots over the window id's
//      bits; waves are then ordered by a scan per window.  So every window's
//      products stay in serial order.  A row's first batch also cuts the row
//      into fold tasks of about kPartTaskProducts products (whole windows).
//   2. k_large_fold (tasks on demand) folds a task's windows one at a time into
//      a dense LDS accumulator of kPartCols sums that start at -0.0
//      (-0.0 + p == p for every p, which is "the first product starts the sum").
//      Each chunk of the window's products is split stably by the low 8 bits of
//      its window column into 256 buckets (one per thread), so a column's
//      products keep their serial order inside its bucket.  When no bucket
//      holds more than kOwnerMax products, thread t adds bucket t's products,
//      in order, to their columns' sums (every column belongs to one thread).
//      Otherwise (hub columns) a second stable split by the high bits sorts
//      the chunk by column, and the first position of each column's run adds
//      the run, left to right, to the column's sum.  No value atomics.  The window's touched columns (a presence bitmap) are
//      then written in column order at the window's place in the structural
//      layout, and only their sums are reset.
//   3. k_large_gapfix moves windows together in the rare rows where a window
//      dropped exact zeros (the stored count of the row goes to annz, and the
//      usual compaction closes the gaps between rows).
#include "device_common.hpp"

namespace spg {
namespace {

using namespace dev;

constexpr uint32_t kPartMask = (uint32_t)(kPartCols - 1);

// Largest bucket (products of one chunk whose window columns share their low 8
// bits) that the fold's owner walk takes; a chunk with a larger one is sorted
// by column and walked by runs, which keeps a hub column's run in a register.
constexpr int kOwnerMax = 32;

// Bits that tell n keys apart.
__device__ __forceinline__ int key_bits(int n) { return n <= 1 ? 0 : 32 - __clz(n - 1); }

// Lanes of the wave holding the same key as this lane, among lanes where
// `valid` (nbits is uniform).
__device__ __forceinline__ uint64_t same_key_lanes(uint32_t key, bool valid, int nbits) {
    uint64_t m = __ballot(valid);
    for (int b = 0; b < nbits; ++b) {
        const bool bit = (key >> b) & 1u;
        const uint64_t bb = __ballot(bit);
        m &= bit ? bb : ~bb;
    }
    return m;
}

// Stable split of the wave's ITEMS items per lane (the wave's serial order:
// item j of lane l is item 64 j + l) by key: rank[j] is the item's position
// among the wave's earlier items of the same key; cnt[key * stride] (zeroed
// by the caller) receives the wave's count per key.  Only this wave touches
// its counters.
template <int ITEMS>
__device__ __forceinline__ void wave_multisplit(const uint32_t (&key)[ITEMS], const bool (&valid)[ITEMS], int nbits,
                                                int *cnt, int stride, int (&rank)[ITEMS], int nitems = ITEMS) {
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        if (j >= nitems) break;  // wave-uniform: items j >= nitems hold nothing
        const uint64_t m = same_key_lanes(key[j], valid[j], nbits);
        const int before = __popcll(m & lt);
        const int base = valid[j] ? cnt[key[j] * stride] : 0;
        rank[j] = base + before;
        wave_sync();  // every lane has read its counter before the group's first lane updates it
        if (valid[j] && before == 0) cnt[key[j] * stride] = base + __popcll(m);
        wave_sync();
    }
}

// Batches of kCapLarge A entries of every large row (the count kernel's
// batches, which the partition kernel takes one at a time).
__global__ void k_part_batches(Operands op, Workspace ws, PartState ps, int count) {
    const RowRec *list = large_list(ws);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const int row = list[i].row;
        ps.segbase[i] = (op.aptr[row + 1] - op.aptr[row] + kCapLarge - 1) / kCapLarge;
    }
}

// Scratch sizes of the large rows: a partitioned row's products, 0 for the
// rows the compact kernel takes.
__global__ void k_part_sizes(Workspace ws, PartState ps, int count) {
    const RowRec *list = large_list(ws);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const int row = list[i].row;
        ps.poff[i] = ws.snnz[row] > ps.skip_le ? ws.ub[row] : 0;
    }
}

template <int TPB, int CAP>
__global__ __launch_bounds__(TPB) void k_large_part(Operands op, Workspace ws, PartState ps, const int64_t *crowptr) {
    static_assert(TPB >= kPartMaxWin, "one thread per window");  // kTpbPart follows kPartMaxWin
    static_assert(CAP == kCapLarge, "a unit of work is one batch of the count kernel");
    constexpr int NW = TPB / 64;
    constexpr int ITEMS = CAP / TPB;
    __shared__ int64_t s_bst[CAP];
    __shared__ double s_av[CAP];
    __shared__ int64_t s_off[CAP + 1];
    __shared__ int64_t s_ws64[NW + 1];
    __shared__ int64_t s_wb[kPartMaxWin];   // next scratch position of each window
    __shared__ int64_t s_np[kPartMaxWin];   // products per window
    __shared__ int64_t s_pex[kPartMaxWin];  // their exclusive scan
    __shared__ int64_t s_dex[kPartMaxWin];  // exclusive scan of the distinct columns per window
    __shared__ int s_cnt[NW][kPartMaxWin];  // per-wave counts, then the waves' offsets inside a chunk
    __shared__ int s_tot[kPartMaxWin];      // a chunk's items per window
    __shared__ int s_gmark[NW][64];         // each wave's entry marks (group_entries)
    __shared__ int s_next;
    const int tid = threadIdx.x, lane = lane_id(), wv = tid >> 6;
    const RowRec *list = large_list(ws);
    const int count = ws.ctrl->bin_count[kBinLarge];
    const int64_t units = (int64_t)ws.ctrl->part_units;
    const int64_t *ubase = ps.segbase + count + 1;  // first batch of each large row
    const int nwin = ps.nwin;
    const int nbits = key_bits(nwin);
    unsigned flags = 0;
    // One unit of work is one batch of kCapLarge A entries of a row: the
    // count kernel recorded each window's products before every batch, so
    // the batches of one row go to their windows' places independently.
    for (int u; (u = next_large_row(&ws.ctrl->part_next, &s_next)) < units;) {
        const int i = upper_entry(ubase, count, (int64_t)u);
        const int64_t bidx = (int64_t)u - ubase[i];
        const int row = list[i].row;
        if (ws.snnz[row] <= ps.skip_le) continue;  // the compact kernel's
        const size_t rb = (size_t)i * (size_t)nwin;
        const int64_t a0 = op.aptr[row], a1 = op.aptr[row + 1];
        const int64_t e0 = a0 + bidx * CAP;
        const bool last = e0 + CAP >= a1;
        const int64_t np = tid < nwin ? ps.pw[rb + tid] : 0;
        const int64_t nd = tid < nwin ? (int64_t)ps.pd[rb + tid] : 0;
        int64_t ptot, dtot;
        const int64_t pex = block_excl_scan<TPB>(np, s_ws64, ptot);
        const int64_t dex = block_excl_scan<TPB>(nd, s_ws64, dtot);
        const int64_t sbase = ps.poff[count + 1 + i];
        const int64_t send = ps.poff[count + 2 + i];
        if (tid < nwin) {
            const size_t seg = (size_t)u * (size_t)nwin + tid;
            s_np[tid] = np;
            s_pex[tid] = pex;
            s_dex[tid] = dex;
            s_wb[tid] = sbase + pex + ps.pseg[seg];
        }
        if (ptot != ws.ub[row] || dtot != ws.snnz[row] || send - sbase != ptot || e0 >= a1) {
            // the counts disagree: nothing of the row is written, and the
            // fold kernel (which reads the flags) folds nothing
            flags |= kErrInternal;
            continue;
        }
        __syncthreads();
        if (tid == 0 && bidx == 0) {
            ws.annz[row] = 0;  // the fold tasks add their stored counts
            // Fold tasks: consecutive windows until a task holds
            // kPartTaskProducts products; empty windows are skipped.
            int nt = 0;
            int64_t acc = 0;
            bool open = false;
            for (int k = 0; k < nwin; ++k) {
                const int64_t n = s_np[k];
                if (n == 0) continue;
                if (!open || acc >= kPartTaskProducts) {
                    ++nt;
                    acc = 0;
                    open = true;
                }
                acc += n;
            }
            const unsigned long long t0 = atomicAdd(&ws.ctrl->ntasks, (unsigned long long)nt);
            if ((int64_t)(t0 + nt) > ps.max_tasks) {
                flags |= kErrInternal;
            } else {
                int64_t t = (int64_t)t0;
                PartTask cur{};
                acc = 0;
                open = false;
                for (int k = 0; k < nwin; ++k) {
                    const int64_t n = s_np[k];
                    if (n == 0) continue;
                    if (!open || acc >= kPartTaskProducts) {
                        if (open) {
                            cur.w1 = (int16_t)k;
                            ps.task[t++] = cur;
                        }
                        cur.s0 = sbase + s_pex[k];
                        cur.out0 = crowptr[row] + s_dex[k];
                        cur.i = i;
                        cur.w0 = (int16_t)k;
                        acc = 0;
                        open = true;
                    }
                    acc += n;
                }
                if (open) {
                    cur.w1 = (int16_t)nwin;
                    ps.task[t++] = cur;
                }
            }
        }
        // The batch's products, in serial order, to their windows.
        const int nbt = (int)min((int64_t)CAP, a1 - e0);
        {
            // each thread a run of consecutive entries: B row start, A value,
            // and the exclusive offset of the entry's products
            const int ipt = (nbt + TPB - 1) / TPB;
            int64_t run = 0;
            for (int j = 0; j < ipt; ++j) {
                const int e = tid * ipt + j;
                if (e < nbt) {
                    const int64_t len = ws.elen[e0 + e];
                    s_bst[e] = entry_bstart_any(ws, e0 + e);
                    s_av[e] = op.aval[e0 + e];
                    s_off[e] = len;
                    run += len;
                }
            }
            int64_t Pb;
            run = block_excl_scan<TPB>(run, s_ws64, Pb);
            for (int j = 0; j < ipt; ++j) {
                const int e = tid * ipt + j;
                if (e < nbt) {
                    const int64_t len = s_off[e];
                    s_off[e] = run;
                    run += len;
                }
            }
            __syncthreads();
            // Chunks of CAP products in serial order, software pipelined: a
            // chunk's B gathers are issued before the previous chunk is split
            // and stored.
            int ncol[ITEMS];
            double nbv[ITEMS], nav[ITEMS];
            bool nok[ITEMS];
            auto issue = [&](int64_t c0) {
                int64_t sx[ITEMS];
                int ent[ITEMS];
                const int64_t g0 = uni64(c0 + (int64_t)wv * (64 * ITEMS));
#pragma unroll
                for (int j = 0; j < ITEMS; ++j) {
                    sx[j] = g0 + j * 64 + lane;
                    nok[j] = sx[j] < Pb;
                }
                group_entries<ITEMS>(s_off, nbt, g0, (int64_t)64, s_gmark[wv], ent);
#pragma unroll
                for (int j = 0; j < ITEMS; ++j) {
                    const int e = ent[j];
                    const int64_t idx = nok[j] ? s_bst[e] + (sx[j] - s_off[e]) : 0;
                    nav[j] = s_av[e];
                    ncol[j] = op.bcol[idx];
                    nbv[j] = op.bval[idx];
                }
            };
            if (Pb > 0) issue(0);
            for (int64_t c0 = 0; c0 < Pb; c0 += CAP) {
                uint32_t key[ITEMS];
                bool ok[ITEMS];
                uint16_t lc[ITEMS];
                double v[ITEMS];
                int rank[ITEMS];
#pragma unroll
                for (int j = 0; j < ITEMS; ++j) {
                    ok[j] = nok[j];
                    const uint32_t c = (uint32_t)ncol[j];
                    v[j] = nav[j] * nbv[j];
                    key[j] = ok[j] ? c >> kPartLog : 0u;
                    lc[j] = (uint16_t)(c & kPartMask);
                    if (ok[j] && (int)key[j] >= nwin) {  // outside B's columns: cannot be counted
                        ok[j] = false;
                        key[j] = 0u;
                        flags |= kErrInternal;
                    }
                }
                if (c0 + CAP < Pb) issue(c0 + CAP);
                for (int q = lane; q < nwin; q += 64) s_cnt[wv][q] = 0;
                wave_sync();
                wave_multisplit<ITEMS>(key, ok, nbits, s_cnt[wv], 1, rank);
                __syncthreads();
                if (tid < nwin) {
                    int r = 0;
#pragma unroll
                    for (int w = 0; w < NW; ++w) {
                        const int c = s_cnt[w][tid];
                        s_cnt[w][tid] = r;
                        r += c;
                    }
                    s_tot[tid] = r;
                }
                __syncthreads();
#pragma unroll
                for (int j = 0; j < ITEMS; ++j) {
                    if (ok[j]) {
                        const int64_t dst = s_wb[key[j]] + s_cnt[wv][key[j]] + rank[j];
                        if (dst < send) {
                            ps.scol[dst] = lc[j];
                            ps.sval[dst] = v[j];
                        } else {
                            flags |= kErrInternal;
                        }
                    }
                }
                __syncthreads();
                if (tid < nwin) s_wb[tid] += s_tot[tid];
                __syncthreads();
            }
        }
        // every window filled exactly up to where the next batch starts
        if (tid < nwin) {
            const int64_t want = sbase + s_pex[tid] +
                                 (last ? s_np[tid] : (int64_t)ps.pseg[((size_t)u + 1) * (size_t)nwin + tid]);
            if (s_wb[tid] != want) flags |= kErrInternal;
        }
        __syncthreads();  // window state is rewritten by the next unit
    }
    if (flags) atomicOr(&ws.ctrl->flags, (unsigned long long)flags);
}

template <int TPB, int CAP>
__global__ __launch_bounds__(TPB) void k_large_fold(Operands op, Workspace ws, PartState ps, int32_t *ccol,
                                                    double *cval) {
    constexpr int NW = TPB / 64;
    constexpr int ITEMS = CAP / TPB;
    constexpr int NB = TPB;                       // buckets of a chunk: one per thread
    constexpr int NBITS = 8;                      // the window column's low bits name the bucket
    constexpr int WORDS = (int)(kPartCols / 32);  // presence bitmap of a window
    static_assert(NB == 1 << NBITS && kPartLog > NBITS && kPartLog <= 16, "one bucket per thread, 16-bit columns");
    static_assert(WORDS <= TPB && kPartCols == 16 * TPB, "extraction: half a bitmap word per thread");
    __shared__ double s_acc[kPartCols];  // running sums, -0.0 where untouched
    __shared__ uint32_t s_bits[WORDS];   // columns touched in the window
    __shared__ uint16_t s_lc[CAP];       // a chunk's products by bucket (then serial order)
    __shared__ __align__(16) double s_lv[CAP];
    __shared__ int s_gc[NB * NW];        // per (bucket, wave) counts, then their offsets; bucket-major
    __shared__ uint64_t s_hm[CAP / 64];  // heads of the sorted chunk's column runs
    __shared__ int s_ws[NW + 1];
    __shared__ int s_mx[NW];
    __shared__ int s_next;
    const int tid = threadIdx.x, lane = lane_id(), wv = tid >> 6;
    const RowRec *list = large_list(ws);
    // Tasks of a partition that reported an inconsistency are not folded.
    const unsigned long long ntk = ws.ctrl->ntasks;
    const int ntasks = (ws.ctrl->flags & kErrInternal) || ntk > (unsigned long long)ps.max_tasks ? 0 : (int)ntk;
    const int nwin = ps.nwin;
    for (int q = tid; q < (int)kPartCols; q += TPB) s_acc[q] = -0.0;
    for (int q = tid; q < WORDS; q += TPB) s_bits[q] = 0u;
    __syncthreads();
    unsigned flags = 0;
    unsigned long long stored = 0;
    for (int t; (t = next_large_row(&ws.ctrl->fold_next, &s_next)) < ntasks;) {
        const PartTask task = ps.task[t];
        const int i = task.i;
        const int row = list[i].row;
        const size_t rb = (size_t)i * (size_t)nwin;
        int64_t s = task.s0, out = task.out0;
        unsigned long long task_stored = 0;
        for (int w = task.w0; w < task.w1; ++w) {
            const int64_t n = ps.pw[rb + w];
            const int d = ps.pd[rb + w];
            if (n == 0) continue;
            // Each chunk's products are loaded one chunk ahead.  A chunk of
            // `total` products gives each lane it = ceil(total / TPB) items
            // (at most ITEMS), item j of lane l of wave wv being product
            // wv * 64 * it + 64 j + l: every wave holds a contiguous range, in
            // order, and a small chunk is spread over all waves.
            uint16_t nlc[ITEMS];
            double nv[ITEMS];
            auto items_of = [&](int64_t c0) {
                const int64_t rem = n - c0;
                return rem >= CAP ? ITEMS : (int)((rem + TPB - 1) / TPB);
            };
            auto load = [&](int64_t c0) {
                const int it = items_of(c0);
                const int64_t tot = min((int64_t)CAP, n - c0);
#pragma unroll
                for (int j = 0; j < ITEMS; ++j) {
                    const int x = wv * (64 * it) + j * 64 + lane;
                    const bool in = j < it && x < tot;
                    nlc[j] = in ? ps.scol[s + c0 + x] : (uint16_t)0;
                    nv[j] = in ? ps.sval[s + c0 + x] : 0.0;
                }
            };
            load(0);
            for (int64_t c0 = 0; c0 < n; c0 += CAP) {
                const int total = (int)min((int64_t)CAP, n - c0);
                const int it = items_of(c0);
                uint16_t lc[ITEMS];
                double v[ITEMS];
                bool ok[ITEMS];
                uint32_t key[ITEMS];
                int rank[ITEMS];
#pragma unroll
                for (int j = 0; j < ITEMS; ++j) {
                    lc[j] = nlc[j];
                    v[j] = nv[j];
                    ok[j] = j < it && wv * (64 * it) + j * 64 + lane < total;
                    key[j] = (uint32_t)lc[j] & (uint32_t)(NB - 1);
                }
                if (c0 + CAP < n) load(c0 + CAP);
                // Stable split by bucket: item j of lane l of wave wv is the
                // chunk's item wv * 64 * it + 64 j + l (serial order); a
                // wave's items of one bucket are ranked by ballots, and the
                // waves are ordered by a scan per bucket.
                for (int k = lane; k < NB; k += 64) s_gc[k * NW + wv] = 0;
                wave_sync();
                wave_multisplit<ITEMS>(key, ok, NBITS, &s_gc[wv], NW, rank, it);
                __syncthreads();
                int cw[NW], cnt = 0;
#pragma unroll
                for (int q = 0; q < NW; ++q) {
                    cw[q] = s_gc[tid * NW + q];
                    cnt += cw[q];
                }
                const int wmx = wave_max_dpp(cnt);
                if (lane == 0) s_mx[wv] = wmx;
                int tot;
                const int st = block_excl_scan<TPB>(cnt, s_ws, tot);  // its barriers also publish s_mx
                int mx = 0;
#pragma unroll
                for (int q = 0; q < NW; ++q) mx = max(mx, s_mx[q]);
                {
                    int o = st;
#pragma unroll
                    for (int q = 0; q < NW; ++q) {
                        s_gc[tid * NW + q] = o;
                        o += cw[q];
                    }
                }
                __syncthreads();
#pragma unroll
                for (int j = 0; j < ITEMS; ++j) {
                    if (ok[j]) {
                        const int dst = s_gc[key[j] * NW + wv] + rank[j];
                        s_lc[dst] = lc[j];
                        s_lv[dst] = v[j];
                    }
                }
                __syncthreads();
                if (mx <= kOwnerMax) {
                    // Thread tid adds bucket tid's products, in serial order,
                    // to their columns' sums (every column of the bucket is
                    // its own), two at a time; a pair of one column is added
                    // in turn.
                    const int en = st + cnt;
                    for (int x = st; x < en; x += 2) {
                        const bool two = x + 1 < en;
                        const int ca = s_lc[x];
                        const double va = s_lv[x];
                        const int cb = two ? (int)s_lc[x + 1] : ca;
                        const double vb = two ? s_lv[x + 1] : 0.0;
                        const double aa = s_acc[ca];
                        if (cb == ca) {
                            s_acc[ca] = two ? (aa + va) + vb : aa + va;
                        } else {
                            const double ab = s_acc[cb];
                            s_acc[ca] = aa + va;
                            s_acc[cb] = ab + vb;
                            atomicOr(&s_bits[cb >> 5], 1u << (cb & 31));
                        }
                        atomicOr(&s_bits[ca >> 5], 1u << (ca & 31));
                    }
                } else {
                    // A bucket holds a hub column's long run: a second stable
                    // split, by the column's high bits, sorts the chunk by
                    // column (position x of the first split is item j of lane
                    // l of wave wv, x = wv * 64 * it + 64 j + l).
                    constexpr int HBITS = kPartLog - NBITS;
#pragma unroll
                    for (int j = 0; j < ITEMS; ++j) {
                        const int x = wv * (64 * it) + j * 64 + lane;
                        lc[j] = ok[j] ? s_lc[x] : (uint16_t)0;
                        v[j] = ok[j] ? s_lv[x] : 0.0;
                        key[j] = (uint32_t)lc[j] >> NBITS;
                    }
                    for (int k = lane; k < (1 << HBITS); k += 64) s_gc[k * NW + wv] = 0;
                    __syncthreads();  // every item is read back before the split overwrites
                    wave_multisplit<ITEMS>(key, ok, HBITS, &s_gc[wv], NW, rank, it);
                    __syncthreads();
                    // offsets of the (1 << HBITS) high buckets, in order, and
                    // of each wave's items inside them
                    {
                        int hc = 0;
                        if (tid < (1 << HBITS)) {
#pragma unroll
                            for (int q = 0; q < NW; ++q) hc += s_gc[tid * NW + q];
                        }
                        int htot;
                        int o = block_excl_scan<TPB>(hc, s_ws, htot);
                        if (tid < (1 << HBITS)) {
#pragma unroll
                            for (int q = 0; q < NW; ++q) {
                                const int c = s_gc[tid * NW + q];
                                s_gc[tid * NW + q] = o;
                                o += c;
                            }
                        }
                    }
                    __syncthreads();
#pragma unroll
                    for (int j = 0; j < ITEMS; ++j) {
                        if (ok[j]) {
                            const int dst = s_gc[key[j] * NW + wv] + rank[j];
                            s_lc[dst] = lc[j];
                            s_lv[dst] = v[j];
                        }
                    }
                    __syncthreads();
                    // Each column's run folds, in serial order, onto its
                    // running sum at the run's first position (x = q * TPB +
                    // tid).  The runs' heads are marked first, one bitmap word
                    // per 64 sorted positions, so that a run's end is the next
                    // head, or the chunk's end, wherever it lies; its values
                    // are then read eight at a time, 16 bytes per read, with
                    // no wait on each next column.
                    int hc[ITEMS];
                    bool hd[ITEMS];
                    uint64_t hm[ITEMS];
#pragma unroll
                    for (int q = 0; q < ITEMS; ++q) {
                        const int x = q * TPB + tid;
                        const bool in = x < total;
                        hc[q] = in ? (int)s_lc[x] : -1;
                        const int pc = in && x > 0 ? (int)s_lc[x - 1] : -2;
                        hd[q] = in && pc != hc[q];
                        hm[q] = __ballot(hd[q]);
                        if (lane == 0) s_hm[x >> 6] = hm[q];
                    }
                    __syncthreads();
#pragma unroll
                    for (int q = 0; q < ITEMS; ++q) {
                        if (hd[q]) {
                            const int x = q * TPB + tid;
                            const int c = hc[q];
                            const uint64_t above = lane == 63 ? 0ull : hm[q] & (~0ull << (lane + 1));
                            int end = total;
                            if (above) {
                                end = x - lane + (__ffsll((long long)above) - 1);
                            } else {
                                for (int k = (x >> 6) + 1; k * 64 < total; ++k) {
                                    const uint64_t h = s_hm[k];
                                    if (h) {
                                        end = k * 64 + (__ffsll((long long)h) - 1);
                                        break;
                                    }
                                }
                            }
                            double acc = s_acc[c];
                            int y = x;
                            if (y & 1) acc = acc + s_lv[y++];
                            for (; y + 8 <= end; y += 8) {
                                const double2 p0 = *reinterpret_cast<const double2 *>(&s_lv[y]);
                                const double2 p1 = *reinterpret_cast<const double2 *>(&s_lv[y + 2]);
                                const double2 p2 = *reinterpret_cast<const double2 *>(&s_lv[y + 4]);
                                const double2 p3 = *reinterpret_cast<const double2 *>(&s_lv[y + 6]);
                                acc = acc + p0.x;
                                acc = acc + p0.y;
                                acc = acc + p1.x;
                                acc = acc + p1.y;
                                acc = acc + p2.x;
                                acc = acc + p2.y;
                                acc = acc + p3.x;
                                acc = acc + p3.y;
                            }
                            for (; y + 2 <= end; y += 2) {
                                const double2 p0 = *reinterpret_cast<const double2 *>(&s_lv[y]);
                                acc = acc + p0.x;
                                acc = acc + p0.y;
                            }
                            if (y < end) acc = acc + s_lv[y];
                            s_acc[c] = acc;
                            atomicOr(&s_bits[c >> 5], 1u << (c & 31));
                        }
                    }
                }
                __syncthreads();
            }
            // The window's sums in column order, exact zeros dropped unless
            // kept.  The touched columns are listed in column order first
            // (thread t reads columns 16t..16t+15 of the bitmap), so that
            // entry i of the list is read by thread i mod TPB and written to
            // C coalesced.  The list takes the chunk's dead value storage.
            uint16_t *s_list = reinterpret_cast<uint16_t *>(s_lv);
            const uint32_t hw = (s_bits[tid >> 1] >> ((tid & 1) * 16)) & 0xffffu;
            int tot;
            int o = block_excl_scan<TPB>((int)__popc(hw), s_ws, tot);
            for (uint32_t r = hw; r; r &= r - 1u) s_list[o++] = (uint16_t)(tid * 16 + __ffs((int)r) - 1);
            __syncthreads();
            if (tid < WORDS) s_bits[tid] = 0u;
            const bool fits = tot == d;  // touched columns are the counted distinct columns
            if (!fits) flags |= kErrInternal;
            const int32_t cbase = (int32_t)((int64_t)w << kPartLog);
            int zeros = 0;
            for (int e = tid; e < tot; e += TPB) {
                const int c = s_list[e];
                const double x = s_acc[c];
                zeros |= !op.keep_zeros && x == 0.0;
                if (fits) {
                    __builtin_nontemporal_store(cbase + c, &ccol[out + e]);
                    __builtin_nontemporal_store(x, &cval[out + e]);
                }
            }
            int st = tot;
            if (__syncthreads_or(zeros) && fits) {
                // Rare: sums of exactly zero.  Rounds of TPB list entries move
                // the nonzero ones down over the window's range (a round reads
                // before it writes, and writes below where later rounds read);
                // the barrier orders them after the stores above.
                __syncthreads();
                st = 0;
                for (int e0 = 0; e0 < tot; e0 += TPB) {
                    const int e = e0 + tid;
                    const double x = e < tot ? s_acc[s_list[e]] : 0.0;
                    const int32_t cc = e < tot ? cbase + (int)s_list[e] : 0;
                    const int nz = e < tot && x != 0.0 ? 1 : 0;
                    int rt;
                    const int ex = block_excl_scan<TPB>(nz, s_ws, rt);
                    if (nz) {
                        ccol[out + st + ex] = cc;
                        cval[out + st + ex] = x;
                    }
                    st += rt;
                }
            }
            for (int e = tid; e < tot; e += TPB) s_acc[s_list[e]] = -0.0;
            if (fits) {
                if (tid == 0) {
                    ps.pst[rb + w] = (int32_t)st;
                    if (st < d) ps.pgap[i] = 1;
                }
                task_stored += (unsigned long long)st;
            }
            out += d;
            s += n;
            __syncthreads();  // the accumulator is clean for the next window
        }
        if (tid == 0 && task_stored)
            atomicAdd(reinterpret_cast<unsigned long long *>(&ws.annz[row]), task_stored);
        stored += task_stored;
    }
    if (tid == 0 && stored) atomicAdd(&ws.ctrl->actual, stored);
    if (flags) atomicOr(&ws.ctrl->flags, (unsigned long long)flags);
}

// Rows where a window dropped exact zeros: the windows' stored runs move
// together to the row's start, in column order (each run moves left, copied
// in ascending blocks, so no entry is overwritten before it is read).
template <int TPB>
__global__ __launch_bounds__(TPB) void k_large_gapfix(Workspace ws, PartState ps, const int64_t *crowptr,
                                                      int32_t *ccol, double *cval) {
    const RowRec *list = large_list(ws);
    const int count = ws.ctrl->bin_count[kBinLarge];
    const int nwin = ps.nwin;
    for (int i = blockIdx.x; i < count; i += gridDim.x) {
        if (!ps.pgap[i]) continue;
        const int row = list[i].row;
        const size_t rb = (size_t)i * (size_t)nwin;
        int64_t dst = crowptr[row], src = crowptr[row];
        for (int w = 0; w < nwin; ++w) {
            const int d = ps.pd[rb + w];
            const int st = ps.pw[rb + w] > 0 ? ps.pst[rb + w] : 0;
            if (dst != src) {
                for (int q0 = 0; q0 < st; q0 += TPB) {
                    const int q = q0 + threadIdx.x;
                    int32_t c = 0;
                    double v = 0.0;
                    if (q < st) {
                        c = ccol[src + q];
                        v = cval[src + q];
                    }
                    __syncthreads();
                    if (q < st) {
                        ccol[dst + q] = c;
                        cval[dst + q] = v;
                    }
                    __syncthreads();
                }
            }
            dst += st;
            src += d;
        }
    }
}

}  // namespace

void launch_part_batches(const Operands &op, const Workspace &ws, const PartState &ps, int nlarge,
                         hipStream_t stream) {
    if (nlarge <= 0) return;
    const int gs = (nlarge + 255) / 256 < 1024 ? (nlarge + 255) / 256 : 1024;
    hipLaunchKernelGGL(k_part_batches, dim3(gs), dim3(256), 0, stream, op, ws, ps, nlarge);
    launch_scan(ps.segbase, ps.segbase + nlarge + 1, nlarge, ws.scan_tmp, &ws.ctrl->part_units, stream);
}

void launch_part(const Operands &op, const Workspace &ws, const PartState &ps, int nlarge, int cus,
                 const int64_t *crowptr, int32_t *ccol, double *cval, hipStream_t stream) {
    if (nlarge <= 0) return;
    launch_zero_bytes(ps.pgap, (size_t)nlarge * 4, stream);
    const int gs = (nlarge + 255) / 256 < 1024 ? (nlarge + 255) / 256 : 1024;
    hipLaunchKernelGGL(k_part_sizes, dim3(gs), dim3(256), 0, stream, ws, ps, nlarge);
    launch_scan(ps.poff, ps.poff + nlarge + 1, nlarge, ws.scan_tmp, nullptr, stream);
    const int gp = cus * 4;
    hipLaunchKernelGGL((k_large_part<kTpbPart, kCapPart>), dim3(gp), dim3(kTpbPart), 0, stream, op, ws, ps, crowptr);
    hipLaunchKernelGGL((k_large_fold<kTpbFold, kCapFold>), dim3(cus * 3), dim3(kTpbFold), 0, stream, op, ws, ps,
                       ccol, cval);
    hipLaunchKernelGGL((k_large_gapfix<256>), dim3(gp), dim3(256), 0, stream, ws, ps, crowptr, ccol, cval);
}

}  // namespace spg
