// fused.hip -- single-pass product of the rows whose products fit one wave.
//
// Semantics are those of the reference's local multiply
// (src/baselines/spgemm_exchange.cpp:274-309): per output row, products a*b in
// serial order (A entries in stored order, then B entries in stored order), a
// column's first product starts its sum and later ones are added left to
// right, columns ascending, sums equal to 0.0 dropped.
//
// A row with at most 256 products is computed entirely by one wave (four
// products per lane): its products are expanded with their serial index s,
// counted into 256 buckets over the row's column range with LDS atomics,
// scattered by bucket, ranked inside their bucket by (column, s) -- the rank is
// the final sorted position -- and folded left to right at each column's first
// position.  Sorted position x lives in slot x / 64 of lane x % 64, so the
// stores of a row are 64 consecutive entries per instruction.  The row's
// stored entries stay in registers until its output offset is known.
//
// Rows are taken in tiles of consecutive rows, one tile per workgroup in
// workgroup order (a global ticket counter would serialise: device-scope
// atomics on one address cost tens of nanoseconds each).  A tile publishes
// its entry count, finds its output offset by looking back over the published
// counts of earlier tiles (decoupled look-back), and writes its rows at their
// final place: no separate counting pass and no scan over the rows.  Rows with
// more products are "reserved": their exact distinct-column count was computed
// before this kernel by the symbolic kernels, and the numeric kernels fill them
// in afterwards.
#include "device_common.hpp"

#include <type_traits>

namespace spg {
namespace {

using namespace dev;

constexpr int kFItems = kFusedCap / 64;   // products per lane
constexpr int kFNb = 256;                 // buckets per row
constexpr int kFKeys = 2 * kFusedCap;     // sorted keys, then padding read by the rank loop
constexpr uint64_t kStAgg = 1ull << 62;   // tile status: count of the tile only
constexpr uint64_t kStIncl = 1ull << 63;  // tile status: count of tiles 0..t
constexpr uint64_t kStVal = kStAgg - 1;
constexpr int kSpinLimit = 1 << 22;       // never reached unless a tile fails to publish
// Wait between polls of an unpublished earlier tile, in units of 64 clocks.
// With one status per line (kStatusStride) the wait length hardly matters
// between 2 and 16; with packed statuses polling contended for one memory
// channel and only the longest wait (127) kept it in check.
#ifndef SPG_LB_SLEEP
#define SPG_LB_SLEEP 8
#endif

// Statuses read per look-back probe (lanes 0 .. kLbWidth-1 of wave 0, one
// line each).  The nearest inclusive total is usually a few tiles back, and a
// narrower probe returns sooner.  On the 2^20-row 16-per-row product with
// 16-row tiles the single-pass kernel took 1.71 ms with 64 statuses per probe,
// 1.57 with 16, 1.54 with 12, 1.57 with 8 and 1.66 with 4; with 8-row tiles
// 1.512 ms with 12, 1.501 with 16, 1.523 with 24 and 1.568 with 8.
#ifndef SPG_LB_WIDTH
#define SPG_LB_WIDTH 16
#endif
constexpr int kLbWidth = SPG_LB_WIDTH;
static_assert(kLbWidth >= 1 && kLbWidth <= 64, "one status per lane");

// A wave's LDS.  The row's entry metadata is dead once its gathers are
// issued, so it shares storage with the sorted products.
template <bool SMALLB>
struct FusedWave {
    using bst_t = typename std::conditional<SMALLB, uint32_t, int64_t>::type;
    // The rows' entry metadata is dead once their gathers are issued (and
    // their items are in registers), so the sort's arrays share its storage.
    union {
        struct {
            double av[kFusedCap];    // A value of each of the row's entries
            bst_t bst[kFusedCap];    // first B entry of each entry's B row
            int off[kFusedCap + 1];  // serial index of each entry's first product
        };
        struct {
            uint32_t key[kFKeys];        // bucket-ordered keys (+ padding), then sorted columns
            double val[kFusedCap + 64];  // products at their sorted positions (+ one dummy per lane)
        };
    };
    alignas(16) int start[kFNb + 4];  // bucket counts, then bucket starts; [kFNb+1] is a dummy counter
    uint8_t sx[kFKeys];               // serial index by bucket position (wide rows)
};

template <bool SMALLB>
struct MetaReg {
    typename FusedWave<SMALLB>::bst_t bst;
    int len;
    double av;
};

// Row results held in registers until the tile's output offset is known:
// slot j of a lane is sorted position lane + 64*j of the row.  Byte j of exnz
// is the slot's stored flag (bit 7) and the count of stored entries of slot j
// in lower lanes (bits 0-5); byte j of tot (uniform) is slot j's stored count,
// so slot j's output position is (stored entries of slots < j) + that count.
struct RowOut {
    int col[kFItems];
    double val[kFItems];
    uint32_t exnz;
    uint32_t tot;
};

template <typename T>
__device__ __forceinline__ T ld_byte(const T *base, uint32_t byte_off) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + byte_off);
}

// Entry lane of a row: lanes past the row's last entry re-read it, so the
// load is unconditional (no register initialisation that would wait on it).
// DIRECT: the B row comes from the entry's column through B's 32-bit row
// offsets (ws.bdir32) instead of the analysis stream; a column outside B's
// rows reads as an empty row and raises kErrRange.  na must be positive.
template <bool SMALLB, bool DIRECT = false>
__device__ __forceinline__ MetaReg<SMALLB> load_meta(const Operands &op, const Workspace &ws, int64_t a0, int na,
                                                     int c0, unsigned &flags) {
    using bst_t = typename FusedWave<SMALLB>::bst_t;
    const int64_t e = a0 + min(c0 + lane_id(), max(na - 1, 0));
    MetaReg<SMALLB> m;
    if constexpr (DIRECT) {
        const int c = __builtin_nontemporal_load(&op.acol[e]);
        const bool ok = (unsigned)c < (unsigned)op.ndir;
        if (__ballot(!ok)) flags |= kErrRange;  // uniform: the kernel publishes lane 0's flags
        const uint32_t b0 = ws.bdir32[ok ? c : 0], b1 = ws.bdir32[ok ? c + 1 : 0];
        m.bst = (bst_t)b0;
        m.len = ok ? (int)(b1 - b0) : 0;
    } else {
        m.bst = (bst_t)entry_bstart<SMALLB>(ws, e);  // small B: 32-bit stream
        m.len = __builtin_nontemporal_load(&ws.elen[e]);
    }
    m.av = __builtin_nontemporal_load(&op.aval[e]);
    return m;
}

// A row's products in registers: slot j of each lane.
struct RowItems {
    int col[kFItems];
    int sidx[kFItems];   // serial index, -1 for an empty slot
    double bv[kFItems];  // B value, then the product
    double av[kFItems];  // A value
};

// Entries a0 .. a0+E-1 (E <= kFusedCap) into LDS: B row start, A value and
// the exclusive prefix of the B row lengths over all E entries (off[E] is
// the total).  m0 holds entries 0..63.  Lanes past E write harmless values
// below kFusedCap; off[E] is written last.
template <bool SMALLB, bool DIRECT = false>
__device__ __forceinline__ void stage_meta(const Operands &op, const Workspace &ws, FusedWave<SMALLB> &w, int64_t a0,
                                           int E, const MetaReg<SMALLB> &m0, unsigned &flags) {
    const int lane = lane_id();
    int carry = 0;
    for (int c0 = 0; c0 < E; c0 += 64) {
        const MetaReg<SMALLB> m = c0 == 0 ? m0 : load_meta<SMALLB, DIRECT>(op, ws, a0, E, c0, flags);
        const int e = c0 + lane;
        const int len = e < E ? m.len : 0;
        const int incl = wave_incl_sum_dpp(len);
        w.bst[e] = m.bst;
        w.av[e] = m.av;
        w.off[e] = carry + incl - len;
        carry += wave_last(incl);
    }
    wave_sync();
    if (lane == 0) w.off[E] = carry;
    wave_sync();
}

// The same over two entry ranges staged one after the other: [aA, aA+nA)
// at positions [0, nA), then [aB, ...) up to E positions in all.
template <bool SMALLB>
__device__ __forceinline__ MetaReg<SMALLB> load_meta2(const Operands &op, const Workspace &ws, int64_t aA, int nA,
                                                      int64_t aB, int E, int c0) {
    using bst_t = typename FusedWave<SMALLB>::bst_t;
    const int p = min(c0 + lane_id(), max(E - 1, 0));
    const int64_t e = p < nA ? aA + p : aB + (p - nA);
    MetaReg<SMALLB> m;
    m.bst = (bst_t)entry_bstart<SMALLB>(ws, e);  // small B: 32-bit stream
    m.len = __builtin_nontemporal_load(&ws.elen[e]);
    m.av = __builtin_nontemporal_load(&op.aval[e]);
    return m;
}
template <bool SMALLB>
__device__ __forceinline__ void stage_meta2(const Operands &op, const Workspace &ws, FusedWave<SMALLB> &w, int64_t aA,
                                            int nA, int64_t aB, int E, const MetaReg<SMALLB> &m0) {
    const int lane = lane_id();
    int carry = 0;
    for (int c0 = 0; c0 < E; c0 += 64) {
        const MetaReg<SMALLB> m = c0 == 0 ? m0 : load_meta2<SMALLB>(op, ws, aA, nA, aB, E, c0);
        const int e = c0 + lane;
        const int len = e < E ? m.len : 0;
        const int incl = wave_incl_sum_dpp(len);
        w.bst[e] = m.bst;
        w.av[e] = m.av;
        w.off[e] = carry + incl - len;
        carry += wave_last(incl);
    }
    wave_sync();
    if (lane == 0) w.off[E] = carry;
    wave_sync();
}

// Expansion of the row whose entries are staged at [m, m+na) with P products:
// slot j of each lane gets one product's column, serial index and values; the
// B gathers are issued here and consumed later.
template <bool SMALLB>
__device__ __forceinline__ void expand_row(const Operands &op, FusedWave<SMALLB> &w, int m, int na, int P,
                                           RowItems &it, unsigned &flags) {
    using bst_t = typename FusedWave<SMALLB>::bst_t;
    const int lane = lane_id();
    const int sbase = uni(w.off[m]);
    if (w.off[m + na] - sbase != P) flags |= kErrInternal;  // P comes from the same lengths
    int L = 0;
    for (int i = 0; i < na; i += 64) {
        const int e = m + min(i + lane, na - 1);
        L = max(L, w.off[e + 1] - w.off[e]);
    }
    L = wave_max_dpp(L);
    int Lp = 1;
    while (Lp < L) Lp <<= 1;
    Lp = uni(Lp);
    const int logL = 31 - __clz(Lp);
    const int G = Lp <= 64 ? (64 >> logL) : 1;
    const bool entry_major = Lp <= 64 && na <= kFItems * G;
    bst_t bs[kFItems];
    int t[kFItems];
    bool ok[kFItems];
    if (entry_major) {
        // Lp lanes per A entry: contiguous reads of each B row, no search.
        const int sub = lane & (Lp - 1);
        const int grp = lane >> logL;
#pragma unroll
        for (int j = 0; j < kFItems; ++j) {
            const int e = j * G + grp;
            const bool in = e < na;
            const int ec = m + (in ? e : 0);
            const int o = w.off[ec];
            ok[j] = in && sub < w.off[ec + 1] - o;
            bs[j] = w.bst[ec];
            it.av[j] = w.av[ec];
            t[j] = sub;
            it.sidx[j] = ok[j] ? o - sbase + sub : -1;
        }
    } else {
        // Serial order dealt out lane by lane; entries found by binary
        // searches in lockstep.
        int sx[kFItems], ex[kFItems];
#pragma unroll
        for (int j = 0; j < kFItems; ++j) {
            ok[j] = lane + 64 * j < P;
            sx[j] = lane + 64 * j + sbase;
        }
        upper_entries<kFItems>(w.off + m, na, sx, ex);
#pragma unroll
        for (int j = 0; j < kFItems; ++j) {
            const int s = lane + 64 * j;
            const int e = m + (ok[j] ? ex[j] : 0);
            bs[j] = w.bst[e];
            it.av[j] = w.av[e];
            t[j] = ok[j] ? s + sbase - w.off[e] : 0;
            it.sidx[j] = ok[j] ? s : -1;
        }
    }
    // Every slot's loads are issued before any product is formed; empty
    // slots read B entry 0 (the row has products, so B is not empty).
    if constexpr (SMALLB) {
        // Byte offsets fit 32 bits: scalar base + vector offset addressing.
#pragma unroll
        for (int j = 0; j < kFItems; ++j) {
            const uint32_t ix = ok[j] ? bs[j] + (uint32_t)t[j] : 0u;
            it.col[j] = ld_byte(op.bcol, ix * 4u);
            it.bv[j] = ld_byte(op.bval, ix * 8u);
        }
    } else {
#pragma unroll
        for (int j = 0; j < kFItems; ++j) {
            const int64_t ix = ok[j] ? bs[j] + t[j] : 0;
            it.col[j] = op.bcol[ix];
            it.bv[j] = op.bval[ix];
        }
    }
}

__device__ __forceinline__ void form_products(RowItems &it) {
#pragma unroll
    for (int j = 0; j < kFItems; ++j) it.bv[j] = it.av[j] * it.bv[j];
}

// Sort and fold a row's P products (1 <= P <= kFusedCap) held in `it` (after
// form_products).  Returns the stored entry count; `distinct` receives the
// distinct columns.
template <bool SMALLB>
__device__ __forceinline__ int sort_fold(FusedWave<SMALLB> &w, int P, const RowItems &it, RowOut &out,
                                         int &distinct, int keep_zeros) {
    const int lane = lane_id();
    const int(&col)[kFItems] = it.col;
    const int(&sidx)[kFItems] = it.sidx;
    const double(&pv)[kFItems] = it.bv;

    // --- buckets over the row's column range.  Keys are offsets from the
    // row's smallest column (unsigned, order preserving).
    int mn = INT_MAX, mx = INT_MIN;
#pragma unroll
    for (int j = 0; j < kFItems; ++j) {
        mn = sidx[j] >= 0 ? min(mn, col[j]) : mn;
        mx = sidx[j] >= 0 ? max(mx, col[j]) : mx;
    }
    mn = wave_min_dpp(mn);
    mx = wave_max_dpp(mx);
    const uint32_t range = (uint32_t)mx - (uint32_t)mn;
    const int lr = range == 0 ? 0 : 32 - __clz(range);
    const int sh = uni(lr > 8 ? lr - 8 : 0);
    const bool packed = lr <= 24;  // offset << 8 | s fits 32 bits
    reinterpret_cast<int4 *>(w.start)[lane] = make_int4(0, 0, 0, 0);
    wave_sync();
    uint32_t rel[kFItems];
    int bk[kFItems], pos[kFItems];
#pragma unroll
    for (int j = 0; j < kFItems; ++j) {
        const bool ok = sidx[j] >= 0;
        rel[j] = ok ? (uint32_t)col[j] - (uint32_t)mn : 0u;
        bk[j] = ok ? (int)(rel[j] >> sh) : kFNb + 1;  // empty slots count into a dummy
        pos[j] = atomicAdd(&w.start[bk[j]], 1);
    }
    wave_sync();
    {
        const int4 c = reinterpret_cast<const int4 *>(w.start)[lane];
        const int s4 = c.x + c.y + c.z + c.w;
        const int incl = wave_incl_sum_dpp(s4);
        const int ex = incl - s4;
        reinterpret_cast<int4 *>(w.start)[lane] = make_int4(ex, ex + c.x, ex + c.x + c.y, ex + c.x + c.y + c.z);
        if (lane == 63) w.start[kFNb] = incl;
    }
    wave_sync();

    // --- scatter keys by bucket; empty slots write at P + lane, which the
    // padding below then overwrites
    uint32_t key[kFItems];
    int st[kFItems], len[kFItems];
#pragma unroll
    for (int j = 0; j < kFItems; ++j) {
        const bool ok = sidx[j] >= 0;
        const int b = ok ? bk[j] : 0;
        st[j] = w.start[b];
        len[j] = ok ? w.start[b + 1] - st[j] : 0;
        key[j] = packed ? ((rel[j] << 8) | (uint32_t)sidx[j]) : rel[j];
        const int d = ok ? st[j] + pos[j] : P + lane;
        w.key[d] = key[j];
        if (!packed) w.sx[d] = (uint8_t)sidx[j];
    }
    int lmax = 0;
#pragma unroll
    for (int j = 0; j < kFItems; ++j) lmax = max(lmax, len[j]);
    lmax = wave_max_dpp(lmax);
    wave_sync();
#pragma unroll
    for (int j = 0; j < kFItems; ++j) w.key[P + lane + 64 * j] = 0xffffffffu;  // padding: above every key
    wave_sync();

    // --- rank inside the bucket by (column, s): the item's sorted position.
    // Keys after the bucket belong to later buckets or the padding and are all
    // larger, so the packed loop reads on without bounds.
    int rank[kFItems];
#pragma unroll
    for (int j = 0; j < kFItems; ++j) rank[j] = 0;
    if (packed) {
        for (int k = 0; k < lmax; ++k) {
#pragma unroll
            for (int j = 0; j < kFItems; ++j) rank[j] += w.key[st[j] + k] < key[j] ? 1 : 0;
        }
    } else {
        for (int k = 0; k < lmax; ++k) {
#pragma unroll
            for (int j = 0; j < kFItems; ++j) {
                const int y = st[j] + min(k, max(len[j] - 1, 0));
                const uint32_t cy = w.key[y];
                const int sy = w.sx[y];
                rank[j] += ((k < len[j]) & ((cy < key[j]) | ((cy == key[j]) & (sy < sidx[j])))) ? 1 : 0;
            }
        }
    }
    wave_sync();
#pragma unroll
    for (int j = 0; j < kFItems; ++j) {
        const int x = sidx[j] >= 0 ? st[j] + rank[j] : P + lane;
        w.key[x] = (uint32_t)col[j];
        w.val[x] = pv[j];
    }
    wave_sync();

    // --- fold each column's run left to right at its first position; sorted
    // position x = lane + 64*j
    int cc[kFItems];
    double vv[kFItems];
    bool head[kFItems], cont[kFItems];
    bool anycont = false;
#pragma unroll
    for (int j = 0; j < kFItems; ++j) {
        cc[j] = (int)w.key[lane + 64 * j];
        vv[j] = w.val[lane + 64 * j];
    }
    // Neighbouring positions' columns by DPP lane shifts; position 64j - 1
    // (lane 0) is lane 63 of slot j - 1, position 64j + 64 (lane 63) lane 0
    // of slot j + 1.
#pragma unroll
    for (int j = 0; j < kFItems; ++j) {
        const int x = lane + 64 * j;
        const int pl = j > 0 ? __builtin_amdgcn_readlane(cc[j > 0 ? j - 1 : 0], 63) : 0;
        const int nl = j + 1 < kFItems ? __builtin_amdgcn_readlane(cc[j + 1 < kFItems ? j + 1 : j], 0) : 0;
        const int pc = dpp_i<0x138, 0xf>(pl, cc[j]);  // wave_shr 1: lane - 1, lane 0 keeps pl
        const int nc = dpp_i<0x130, 0xf>(nl, cc[j]);  // wave_shl 1: lane + 1, lane 63 keeps nl
        const bool valid = x < P;
        head[j] = valid && (x == 0 || pc != cc[j]);
        cont[j] = x + 1 < P && nc == cc[j];
        anycont |= head[j] & cont[j];
    }
    if (__ballot(anycont)) {
#pragma unroll
        for (int j = 0; j < kFItems; ++j) {
            if (head[j] && cont[j]) {
                double v = vv[j];
                for (int y = lane + 64 * j + 1; y < P && (int)w.key[y] == cc[j]; ++y) v = v + w.val[y];
                vv[j] = v;
            }
        }
    }
    // Output positions: one scan of the four slots' flags packed in bytes
    // (no field exceeds 64, so there is no carry between fields).
    uint32_t packed_nz = 0, packed_head = 0;
#pragma unroll
    for (int j = 0; j < kFItems; ++j) {
        packed_nz |= (head[j] && (keep_zeros || vv[j] != 0.0) ? 1u : 0u) << (8 * j);
        packed_head |= (head[j] ? 1u : 0u) << (8 * j);
    }
    const uint32_t incl = (uint32_t)wave_incl_sum_dpp((int)packed_nz);
    const uint32_t tot = (uint32_t)wave_last((int)incl);
    const uint32_t ex = incl - packed_nz;
    uint32_t before = 0;
#pragma unroll
    for (int j = 0; j < kFItems; ++j) {
        out.col[j] = cc[j];
        out.val[j] = vv[j];
        before += (tot >> (8 * j)) & 0xffu;
    }
    out.exnz = ex | (packed_nz << 7);
    out.tot = tot;
    const uint32_t hs = (uint32_t)wave_last(wave_incl_sum_dpp((int)packed_head));
    distinct = (int)((hs & 0xffu) + ((hs >> 8) & 0xffu) + ((hs >> 16) & 0xffu) + (hs >> 24));
    wave_sync();  // LDS of this wave is reused by its next row
    return (int)before;
}

__device__ __forceinline__ uint64_t ld_status(const uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_status(uint64_t *p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive entry offset of tile t (one wave): publish the tile's count, then
// sum the counts of earlier tiles back to the nearest published inclusive
// total, kLbWidth tiles per probe.
__device__ int64_t tile_lookback(uint64_t *status, int64_t t, int64_t agg, unsigned &flags) {
    const int lane = lane_id();
    if (lane == 0) st_status(&status[t * kStatusStride], (t == 0 ? kStIncl : kStAgg) | (uint64_t)agg);
    if (t == 0) return 0;
    int64_t excl = 0;
    int64_t p = t - 1;
    for (int spins = 0;;) {
        // Lane l < kLbWidth reads the status at distance l behind p; the
        // other lanes read nothing and stand for "not yet known", so a probe
        // ends at distance kLbWidth.
        const int64_t q = p - lane;
        const uint64_t s = lane >= kLbWidth ? 0ull : q >= 0 ? ld_status(&status[q * kStatusStride]) : kStIncl;  // before tile 0: total 0
        // fn: distance of the nearest unpublished tile, fi: of the nearest
        // inclusive total.
        const uint64_t nr_m = __ballot((s & (kStIncl | kStAgg)) == 0);
        const uint64_t incl_m = __ballot((s & kStIncl) != 0);
        const int fn = nr_m ? __ffsll((long long)nr_m) - 1 : 64;
        const int fi = incl_m ? __ffsll((long long)incl_m) - 1 : 64;
        const int lim = fi < fn ? fi + 1 : fn;  // statuses summed: distances < lim
        if (lim > 0) excl += wave_sum(lane < lim ? (int64_t)(s & kStVal) : (int64_t)0);
        if (fi < fn) break;
        if (fn > 0) {
            p -= fn;
        } else {
            // Tile p is unpublished: wait on its status alone (every lane
            // loads the same line) before the next probe.
            bool ready = false;
            while (!ready && ++spins <= kSpinLimit) {
                __builtin_amdgcn_s_sleep(SPG_LB_SLEEP);
                const uint64_t sp = ld_status(&status[p * kStatusStride]);
                ready = (__builtin_amdgcn_readfirstlane((int)(uint32_t)(sp >> 32)) & (int)((kStIncl | kStAgg) >> 32)) != 0;
            }
            if (!ready) {
                flags |= kErrInternal;
                break;
            }
        }
    }
    if (lane == 0) st_status(&status[t * kStatusStride], kStIncl | (uint64_t)(excl + agg));
    return excl;
}

__device__ __forceinline__ int64_t lane64(int64_t v, int k) {
    return (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)v >> 32), k) << 32) |
                     (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, k));
}

// Waves per SIMD the classic tiles' kernel is compiled for (0: no
// constraint); five fit its LDS and 96 VGPRs without spills.
#ifndef SPG_FUSED_WPE
#define SPG_FUSED_WPE 5
#endif
#if SPG_FUSED_WPE
#define SPG_FUSED_ATTR __attribute__((amdgpu_waves_per_eu(SPG_FUSED_WPE)))
#else
#define SPG_FUSED_ATTR
#endif

// DIRECT: every row has at most kFusedCap products and A entries (the host
// checked the longest rows of A and B), so no analysis ran: each row's
// metadata comes from its columns and B's 32-bit row offsets, its product
// count from the staged lengths, and the tile also publishes the statistics
// the analysis would have gathered (products, rows per bin).
template <bool SMALLB, int R, bool DIRECT = false>
__global__ __launch_bounds__(kFusedTileTpb) SPG_FUSED_ATTR void k_fused(Operands op, Workspace ws, FusedOut fo) {
    constexpr int NW = kFusedTileTpb / 64;
    constexpr int NS = kDirectFields - 1;  // DIRECT statistics besides the distinct columns
    __shared__ FusedWave<SMALLB> s_w[NW];
    __shared__ int64_t s_wtot[NW];
    __shared__ int64_t s_base;
    const int wv = threadIdx.x >> 6;
    const int lane = lane_id();
    FusedWave<SMALLB> &w = s_w[wv];
    __shared__ int64_t s_dist[NW];
    __shared__ int64_t s_dstat[DIRECT ? NW : 1][NS];
    unsigned flags = 0;
    int64_t distinct_sum = 0;

    // One tile per workgroup, in workgroup order: every earlier tile belongs
    // to a workgroup dispatched before this one, so the look-back below only
    // waits on workgroups that are running or done.
    const int64_t t = (int64_t)blockIdx.x;
    const int64_t r0 = t * (int64_t)(NW * R) + (int64_t)wv * R;
    // Row offsets and work bounds of this wave's rows (lane k: row r0+k).
    const int64_t ap = lane <= R ? op.aptr[min(r0 + lane, (int64_t)op.arows)] : 0;
    const int64_t rq = min(r0 + lane, (int64_t)op.arows - 1);
    const int64_t ubl = !DIRECT && lane < R ? ws.ub[rq] : 0;
    const int64_t snz = !DIRECT && lane < R ? ws.snnz[rq] : 0;
    RowOut out[R];
    int cnt[R], P[R];
    bool fused[R];
    int64_t wave_total = 0;
    bool joint = true;  // every row single-pass or without products
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int64_t na = lane64(ap, k + 1) - lane64(ap, k);
        if constexpr (DIRECT) {
            fused[k] = r0 + k < op.arows && na > 0;  // products are counted once staged
            P[k] = 0;
        } else {
            const int64_t ub = lane64(ubl, k);
            fused[k] = r0 + k < op.arows && ub > 0 && ub <= kFusedCap && na <= kFusedCap;
            joint &= fused[k] || ub == 0 || r0 + k >= op.arows;
            P[k] = (int)ub;
            if (!fused[k] && r0 + k < op.arows && ub > 0) wave_total += lane64(snz, k);  // reserved for the numeric kernels
        }
        cnt[k] = 0;
        out[k].exnz = 0;
        out[k].tot = 0;
    }
    const int64_t E = lane64(ap, R) - lane64(ap, 0);
    joint &= E <= kFusedCap;
    // First 64 entries of the wave's rows: row 0's metadata on either path.
    MetaReg<SMALLB> meta0;
    if (E > 0 && (fused[0] || joint))
        meta0 = load_meta<SMALLB, DIRECT>(op, ws, lane64(ap, 0), (int)min(E, (int64_t)kFusedCap), 0, flags);
    if (joint) {
        // All rows' metadata staged at once and all rows' gathers in flight
        // together: one memory round trip for the rows of this wave.
        if (E > 0) {
            stage_meta<SMALLB, DIRECT>(op, ws, w, lane64(ap, 0), (int)E, meta0, flags);
            RowItems it[R];
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const int m = (int)(lane64(ap, k) - lane64(ap, 0)), na = (int)(lane64(ap, k + 1) - lane64(ap, k));
                if constexpr (DIRECT) {
                    if (fused[k]) P[k] = uni(w.off[m + na] - w.off[m]);
                    fused[k] = fused[k] && P[k] > 0;
                }
                if (fused[k]) expand_row<SMALLB>(op, w, m, na, P[k], it[k], flags);
            }
#pragma unroll
            for (int k = 0; k < R; ++k)
                if (fused[k]) form_products(it[k]);
            wave_sync();  // metadata reads are complete before the sort reuses its storage
#pragma unroll
            for (int k = 0; k < R; ++k) {
                if (fused[k]) {
                    int dist = 0;
                    cnt[k] = sort_fold<SMALLB>(w, P[k], it[k], out[k], dist, op.keep_zeros);
                    distinct_sum += dist;
                    wave_total += cnt[k];
                }
            }
        }
    } else {
        // Row by row; the next row's metadata is loaded during this one.
        MetaReg<SMALLB> meta[R];
        meta[0] = meta0;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            if (k + 1 < R && fused[k + 1])
                meta[k + 1] = load_meta<SMALLB, DIRECT>(op, ws, lane64(ap, k + 1),
                                                        (int)(lane64(ap, k + 2) - lane64(ap, k + 1)), 0, flags);
            if (fused[k]) {
                const int na = (int)(lane64(ap, k + 1) - lane64(ap, k));
                stage_meta<SMALLB, DIRECT>(op, ws, w, lane64(ap, k), na, meta[k], flags);
                if constexpr (DIRECT) {
                    P[k] = uni(w.off[na] - w.off[0]);
                    fused[k] = P[k] > 0;
                }
                if (fused[k]) {
                    RowItems it;
                    expand_row<SMALLB>(op, w, 0, na, P[k], it, flags);
                    form_products(it);
                    wave_sync();
                    int dist = 0;
                    cnt[k] = sort_fold<SMALLB>(w, P[k], it, out[k], dist, op.keep_zeros);
                    distinct_sum += dist;
                    wave_total += cnt[k];
                } else {
                    wave_sync();  // the staged metadata is dead before the next row stages
                }
            }
        }
    }
    if (lane == 0) {
        s_wtot[wv] = wave_total;
        s_dist[wv] = distinct_sum;
        if constexpr (DIRECT) {
            int64_t st[NS] = {0, 0, 0, 0};  // products, rows without products, rows of W1, W2
#pragma unroll
            for (int k = 0; k < R; ++k) {
                if (r0 + k >= op.arows) continue;
                const int64_t na = lane64(ap, k + 1) - lane64(ap, k);
                st[0] += P[k];
                st[P[k] == 0 ? 1 : (P[k] <= kCapW1 && na <= kCapW1 ? 2 : 3)] += 1;
            }
#pragma unroll
            for (int f = 0; f < NS; ++f) s_dstat[wv][f] = st[f];
        }
    }
    __syncthreads();
    if (wv == 0) {
        int64_t agg = 0, dist = 0;
#pragma unroll
        for (int i = 0; i < NW; ++i) {
            agg += s_wtot[i];
            dist += s_dist[i];
        }
        if (lane == 0) {
            fo.distinct[t] = dist;
            if constexpr (DIRECT) {
#pragma unroll
                for (int f = 0; f < NS; ++f) {
                    int64_t v = 0;
#pragma unroll
                    for (int i = 0; i < NW; ++i) v += s_dstat[i][f];
                    fo.distinct[(int64_t)(f + 1) * fo.ntiles + t] = v;
                }
            }
        }
        const int64_t excl = tile_lookback(fo.status, t, agg, flags);
        if (lane == 0) {
            s_base = excl;
            if (t == fo.ntiles - 1) {
                fo.rowptr[op.arows] = excl + agg;
                ws.ctrl->reserved = (unsigned long long)(excl + agg);
            }
        }
    }
    __syncthreads();
    int64_t base = s_base;
    for (int i = 0; i < wv; ++i) base += s_wtot[i];
    base = uni64(base);
    // Row offsets and stored counts of the wave's rows (lane k: row r0+k).
    int64_t rp = base;
    int64_t my_rp = 0, my_cnt = 0;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        if (lane == k) {
            my_rp = rp;
            my_cnt = cnt[k];
        }
        if (fused[k]) {
            int32_t *cc = fo.ccol + rp;
            double *cv = fo.cval + rp;
            uint32_t before = 0;
#pragma unroll
            for (int j = 0; j < kFItems; ++j) {
                const uint32_t b = (out[k].exnz >> (8 * j)) & 0xffu;
                if (b & 0x80u) {
                    const uint32_t o = before + (b & 0x7fu);
                    __builtin_nontemporal_store(
                        out[k].col[j], reinterpret_cast<int32_t *>(reinterpret_cast<char *>(cc) + o * 4u));
                    __builtin_nontemporal_store(
                        out[k].val[j], reinterpret_cast<double *>(reinterpret_cast<char *>(cv) + o * 8u));
                }
                before += (out[k].tot >> (8 * j)) & 0xffu;
            }
            rp += cnt[k];
        } else if (!DIRECT && r0 + k < op.arows && lane64(ubl, k) > 0) {
            rp += lane64(snz, k);
        }
    }
    if (lane < R && r0 + lane < op.arows) {
        fo.rowptr[r0 + lane] = my_rp;
        bool fl = DIRECT;
#pragma unroll
        for (int k = 0; k < R; ++k) fl |= (lane == k) & fused[k];
        if (fl) ws.annz[r0 + lane] = my_cnt;
    }
    if (lane == 0 && flags) atomicOr(&ws.ctrl->flags, (unsigned long long)flags);
}

// Dealt tiles: 32 consecutive rows per 1024-thread workgroup, whose
// single-pass rows are dealt to the 16 waves in row order, two at most per
// wave.  Every other row's count (reserved, or empty) is folded into the
// tile's total, so the look-back and the row offsets are those of
// consecutive-row tiles.
template <bool SMALLB>
__global__ __launch_bounds__(kFusedDealTpb) void k_fused_dealt(Operands op, Workspace ws, FusedOut fo) {
    constexpr int NW = kFusedDealTpb / 64;
    constexpr int R = 2;                // single-pass rows per wave at most
    constexpr int TR = kFusedDealTile;  // rows per tile
    static_assert(TR <= 64 && NW * R >= TR, "every row of a tile fits its waves");
    __shared__ FusedWave<SMALLB> s_w[NW];
    __shared__ int64_t s_rcnt[TR];  // entries of each tile row
    __shared__ int64_t s_rp[TR];    // output offset of each tile row
    __shared__ int64_t s_dist[NW];
    const int wv = threadIdx.x >> 6;
    const int lane = lane_id();
    FusedWave<SMALLB> &w = s_w[wv];
    unsigned flags = 0;
    int64_t distinct_sum = 0;

    const int64_t t = (int64_t)blockIdx.x;  // tiles in workgroup order, as k_fused
    const int64_t tr0 = t * (int64_t)TR;
    // The tile's rows on lanes 0..TR-1: row offsets, work bounds, counts.
    const bool lrow_ok = lane < TR && tr0 + lane < op.arows;
    const int64_t rq = min(tr0 + lane, (int64_t)op.arows - 1);
    const int64_t ap = lane < TR ? op.aptr[min(tr0 + lane, (int64_t)op.arows)] : 0;
    const int64_t ap1 = lane < TR ? op.aptr[min(tr0 + lane + 1, (int64_t)op.arows)] : 0;
    const int64_t ubl = lrow_ok ? ws.ub[rq] : 0;
    const int64_t snz = lrow_ok ? ws.snnz[rq] : 0;
    // Single-pass rows by the bins' rule (bins W1, W2).
    const uint64_t fmask = __ballot(lrow_ok && ubl > 0 && ubl <= kFusedCap && ap1 - ap <= kFusedCap);
    const int nf = __popcll(fmask);  // <= TR <= NW * R
    // Ranks to waves in row order: with at most NW rows one each, else the
    // first nf - NW waves take two adjacent ranks.
    int first, nmine;
    if (nf <= NW) {
        first = wv;
        nmine = wv < nf ? 1 : 0;
    } else {
        const int extra = nf - NW;
        first = wv + min(wv, extra);
        nmine = wv < extra ? 2 : 1;
    }
    int lrow[R];
    bool fused[R];
    int64_t a0[R];
    int na[R], ubk[R], cnt[R];
    RowOut out[R];
    {
        uint64_t m = fmask;
        for (int i = 0; i < first; ++i) m &= m - 1ull;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            fused[k] = k < nmine;
            lrow[k] = fused[k] ? uni(__ffsll((long long)m) - 1) : 0;
            if (fused[k]) m &= m - 1ull;
            a0[k] = lane64(ap, lrow[k]);
            na[k] = (int)(lane64(ap1, lrow[k]) - a0[k]);
            ubk[k] = (int)lane64(ubl, lrow[k]);
            cnt[k] = 0;
            out[k].exnz = 0;
            out[k].tot = 0;
        }
    }
    // The wave's rows (adjacent or not) with at most kFusedCap entries
    // together are staged one after the other and gathered at once: one
    // memory round trip for both.
    const int nB = nmine == 2 ? na[1] : 0;
    const int E = na[0] + nB;
    const bool joint = fused[0] && E <= kFusedCap;
    MetaReg<SMALLB> meta0;
    if (fused[0]) {
        if (joint) meta0 = load_meta2<SMALLB>(op, ws, a0[0], na[0], a0[1], E, 0);
        else meta0 = load_meta<SMALLB>(op, ws, a0[0], na[0], 0, flags);
    }
    if (joint) {
        stage_meta2<SMALLB>(op, ws, w, a0[0], na[0], a0[1], E, meta0);
        RowItems it[R];
#pragma unroll
        for (int k = 0; k < R; ++k)
            if (fused[k]) expand_row<SMALLB>(op, w, k == 0 ? 0 : na[0], na[k], ubk[k], it[k], flags);
#pragma unroll
        for (int k = 0; k < R; ++k)
            if (fused[k]) form_products(it[k]);
        wave_sync();  // metadata reads are complete before the sort reuses its storage
#pragma unroll
        for (int k = 0; k < R; ++k) {
            if (fused[k]) {
                int dist = 0;
                cnt[k] = sort_fold<SMALLB>(w, ubk[k], it[k], out[k], dist, op.keep_zeros);
                distinct_sum += dist;
            }
        }
    } else if (fused[0]) {
        MetaReg<SMALLB> meta[R];
        meta[0] = meta0;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            if (k + 1 < R && fused[k + 1]) meta[k + 1] = load_meta<SMALLB>(op, ws, a0[k + 1], na[k + 1], 0, flags);
            if (fused[k]) {
                stage_meta<SMALLB>(op, ws, w, a0[k], na[k], meta[k], flags);
                RowItems it;
                expand_row<SMALLB>(op, w, 0, na[k], ubk[k], it, flags);
                form_products(it);
                wave_sync();
                int dist = 0;
                cnt[k] = sort_fold<SMALLB>(w, ubk[k], it, out[k], dist, op.keep_zeros);
                distinct_sum += dist;
            }
        }
    }
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < R; ++k)
            if (fused[k]) s_rcnt[lrow[k]] = cnt[k];
        s_dist[wv] = distinct_sum;
    }
    if (wv == 0 && lane < TR && !((fmask >> lane) & 1ull)) s_rcnt[lane] = ubl > 0 ? snz : 0;  // reserved or empty
    __syncthreads();
    if (wv == 0) {
        const int64_t c = lane < TR ? s_rcnt[lane] : 0;
        const int64_t inc = wave_incl_scan(c);
        const int64_t agg = lane64(inc, TR - 1);
        int64_t dist = 0;
#pragma unroll
        for (int i = 0; i < NW; ++i) dist += s_dist[i];
        if (lane == 0) fo.distinct[t] = dist;
        const int64_t excl = tile_lookback(fo.status, t, agg, flags);
        const int64_t rp = excl + inc - c;
        if (lane < TR) {
            s_rp[lane] = rp;
            if (lrow_ok) {
                fo.rowptr[tr0 + lane] = rp;
                if ((fmask >> lane) & 1ull) ws.annz[tr0 + lane] = c;
            }
        }
        if (lane == 0 && t == fo.ntiles - 1) {
            fo.rowptr[op.arows] = excl + agg;
            ws.ctrl->reserved = (unsigned long long)(excl + agg);
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < R; ++k) {
        if (fused[k]) {
            const int64_t rp = uni64(s_rp[lrow[k]]);
            int32_t *cc = fo.ccol + rp;
            double *cv = fo.cval + rp;
            uint32_t before = 0;
#pragma unroll
            for (int j = 0; j < kFItems; ++j) {
                const uint32_t b = (out[k].exnz >> (8 * j)) & 0xffu;
                if (b & 0x80u) {
                    const uint32_t o = before + (b & 0x7fu);
                    __builtin_nontemporal_store(
                        out[k].col[j], reinterpret_cast<int32_t *>(reinterpret_cast<char *>(cc) + o * 4u));
                    __builtin_nontemporal_store(
                        out[k].val[j], reinterpret_cast<double *>(reinterpret_cast<char *>(cv) + o * 8u));
                }
                before += (out[k].tot >> (8 * j)) & 0xffu;
            }
        }
    }
    if (lane == 0 && flags) atomicOr(&ws.ctrl->flags, (unsigned long long)flags);
}

// Packed tiles per analyze block: max(1, ceil(single-pass rows / cap)).
__global__ void k_pack_counts(const int64_t *part, int64_t nblocks, int64_t *tcnt) {
    for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < nblocks;
         b += (int64_t)gridDim.x * blockDim.x) {
        const int64_t f = part[b * kAnalyzeFields + 2 + kBinW1] + part[b * kAnalyzeFields + 2 + kBinW2];
        tcnt[b] = max((int64_t)1, (f + kFusedPackCap - 1) / kFusedPackCap);
    }
}

// Tile map: tile toff[b] + j is tile j of block b.
__global__ void k_pack_map(const int64_t *tcnt, const int64_t *toff, int64_t nblocks, int64_t *tmap) {
    for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < nblocks;
         b += (int64_t)gridDim.x * blockDim.x) {
        const int64_t o = toff[b];
        for (int64_t j = 0; j < tcnt[b]; ++j) tmap[o + j] = (b << 5) | j;
    }
}

// Packed tiles: tile j of an analyze block computes the block's single-pass
// rows of rank 16j .. 16j+15 (in row order), dealt two per wave, and owns
// the block's rows from the one after the previous tile's last single-pass
// row up to its own last one (the block's last tile: up to the block's end).
// The counts of the other rows in that range (reserved: counted ahead; empty:
// zero) join the tile's total, and the tile writes the row offsets of every
// row it owns.  Tiles run in map order, which is row order, so the look-back
// is that of consecutive-row tiles.
template <bool SMALLB>
__global__ __launch_bounds__(kFusedTpb) void k_fused_packed(Operands op, Workspace ws, FusedOut fo) {
    constexpr int NW = kFusedTpb / 64;
    constexpr int R = 2;
    constexpr int BR = kAnalyzeRows;  // rows per block
    static_assert(NW * R == kFusedPackCap && BR == 256 && NW >= BR / 64, "tile shape");
    __shared__ FusedWave<SMALLB> s_w[NW];
    __shared__ int64_t s_rc[BR];  // count of each owned row of the block (0 for others)
    __shared__ int64_t s_rp[BR];  // output offset of each owned row
    __shared__ uint64_t s_fm[BR / 64];
    __shared__ int64_t s_dist[NW];
    __shared__ int64_t s_a0[BR];  // the block's rows: first A entry, entries, work bound
    __shared__ int s_na[BR], s_ub[BR];
    const int wv = threadIdx.x >> 6;
    const int lane = lane_id();
    const int tid = threadIdx.x;
    FusedWave<SMALLB> &w = s_w[wv];
    unsigned flags = 0;
    int64_t distinct_sum = 0;

    const int64_t t = (int64_t)blockIdx.x;  // tiles in workgroup order
    const int64_t ntiles = (int64_t)ws.ctrl->packed_tiles;
    if (t >= ntiles) {  // launched for an upper bound of the tile count
        if (tid == 0) fo.distinct[t] = 0;
        return;
    }
    const int64_t code = fo.tmap[t];
    const int64_t br0 = (code >> 5) * BR;
    const int jt = (int)(code & 31);
    // The block's rows: thread tid < BR holds row br0 + tid.
    const int64_t r = br0 + tid;
    const bool valid = tid < BR && r < op.arows;
    const int64_t ubr = valid ? ws.ub[r] : 0;
    const int64_t a0r = valid ? op.aptr[r] : 0;
    const int64_t nar = valid ? op.aptr[r + 1] - a0r : 0;
    const bool sp = valid && ubr > 0 && ubr <= kFusedCap && nar <= kFusedCap;
    const int64_t snr = valid && ubr > 0 && !sp ? ws.snnz[r] : 0;
    {
        const uint64_t m = __ballot(sp);
        if (wv < BR / 64 && lane == 0) s_fm[wv] = m;
        if (sp) {  // the single-pass rows' data for the waves that take them
            s_a0[tid] = a0r;
            s_na[tid] = (int)nar;
            s_ub[tid] = (int)ubr;
        }
    }
    __syncthreads();
    uint64_t fm[BR / 64];
    int nfb = 0;
#pragma unroll
    for (int i = 0; i < BR / 64; ++i) {
        fm[i] = s_fm[i];
        nfb += __popcll(fm[i]);
    }
    // Row (offset in the block) of single-pass rank q, -1 if none: the word
    // holding it, then its bit by halving (popcount of the lower half).
    auto rank_row = [&](int q) -> int {
        if (q < 0 || q >= nfb) return -1;
#pragma unroll
        for (int i = 0; i < BR / 64; ++i) {
            const int c = __popcll(fm[i]);
            if (q < c) {
                uint64_t m = fm[i];
                int pos = 0;
#pragma unroll
                for (int width = 32; width > 0; width >>= 1) {
                    const uint64_t low = m & ((1ull << width) - 1ull);
                    const int lc = __popcll(low);
                    if (q >= lc) {
                        q -= lc;
                        m >>= width;
                        pos += width;
                    } else {
                        m = low;
                    }
                }
                return i * 64 + pos;
            }
            q -= c;
        }
        return -1;
    };
    const int q0 = jt * kFusedPackCap;
    const bool last_tile = q0 + kFusedPackCap >= nfb;
    const int lo = jt == 0 ? 0 : rank_row(q0 - 1) + 1;
    const int hi = last_tile ? BR - 1 : rank_row(q0 + kFusedPackCap - 1);
    // This wave's ranks: q0 + 2wv, q0 + 2wv + 1.
    int brow[R];
    bool fused[R];
    int64_t a0[R];
    int na[R], ubk[R], cnt[R];
    RowOut out[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
        brow[k] = rank_row(q0 + R * wv + k);
        fused[k] = brow[k] >= 0;
        a0[k] = fused[k] ? s_a0[brow[k]] : 0;
        na[k] = fused[k] ? s_na[brow[k]] : 0;
        ubk[k] = fused[k] ? s_ub[brow[k]] : 0;
        cnt[k] = 0;
        out[k].exnz = 0;
        out[k].tot = 0;
    }
    const int nB = fused[1] ? na[1] : 0;
    const int E = na[0] + nB;
    const bool joint = fused[0] && E <= kFusedCap;
    MetaReg<SMALLB> meta0;
    if (fused[0]) {
        if (joint) meta0 = load_meta2<SMALLB>(op, ws, a0[0], na[0], a0[1], E, 0);
        else meta0 = load_meta<SMALLB>(op, ws, a0[0], na[0], 0, flags);
    }
    if (joint) {
        stage_meta2<SMALLB>(op, ws, w, a0[0], na[0], a0[1], E, meta0);
        RowItems it[R];
#pragma unroll
        for (int k = 0; k < R; ++k)
            if (fused[k]) expand_row<SMALLB>(op, w, k == 0 ? 0 : na[0], na[k], ubk[k], it[k], flags);
#pragma unroll
        for (int k = 0; k < R; ++k)
            if (fused[k]) form_products(it[k]);
        wave_sync();
#pragma unroll
        for (int k = 0; k < R; ++k) {
            if (fused[k]) {
                int dist = 0;
                cnt[k] = sort_fold<SMALLB>(w, ubk[k], it[k], out[k], dist, op.keep_zeros);
                distinct_sum += dist;
            }
        }
    } else if (fused[0]) {
        MetaReg<SMALLB> meta[R];
        meta[0] = meta0;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            if (k + 1 < R && fused[k + 1]) meta[k + 1] = load_meta<SMALLB>(op, ws, a0[k + 1], na[k + 1], 0, flags);
            if (fused[k]) {
                stage_meta<SMALLB>(op, ws, w, a0[k], na[k], meta[k], flags);
                RowItems it;
                expand_row<SMALLB>(op, w, 0, na[k], ubk[k], it, flags);
                form_products(it);
                wave_sync();
                int dist = 0;
                cnt[k] = sort_fold<SMALLB>(w, ubk[k], it, out[k], dist, op.keep_zeros);
                distinct_sum += dist;
            }
        }
    }
    // Counts of the owned rows: the single-pass rows of this tile from their
    // waves, the others from the count kernels (or zero).
    if (tid < BR) s_rc[tid] = (tid >= lo && tid <= hi && !sp) ? snr : 0;
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < R; ++k)
            if (fused[k]) s_rc[brow[k]] = cnt[k];
        s_dist[wv] = distinct_sum;
    }
    __syncthreads();
    if (wv == 0) {
        // Scan of the block's BR counts, 4 consecutive rows per lane.
        int64_t c[BR / 64];
        int64_t loc = 0;
#pragma unroll
        for (int k = 0; k < BR / 64; ++k) {
            c[k] = s_rc[lane * (BR / 64) + k];
            loc += c[k];
        }
        const int64_t inc = wave_incl_scan(loc);
        const int64_t agg = lane64(inc, 63);
        int64_t dist = 0;
#pragma unroll
        for (int i = 0; i < NW; ++i) dist += s_dist[i];
        if (lane == 0) fo.distinct[t] = dist;
        const int64_t excl = tile_lookback(fo.status, t, agg, flags);
        int64_t run = excl + inc - loc;
#pragma unroll
        for (int k = 0; k < BR / 64; ++k) {
            const int o = lane * (BR / 64) + k;
            s_rp[o] = run;
            const int64_t rr = br0 + o;
            if (o >= lo && o <= hi && rr < op.arows) {
                fo.rowptr[rr] = run;
                if ((fm[o >> 6] >> (o & 63)) & 1ull) ws.annz[rr] = c[k];
            }
            run += c[k];
        }
        if (lane == 0 && t == ntiles - 1) {
            fo.rowptr[op.arows] = excl + agg;
            ws.ctrl->reserved = (unsigned long long)(excl + agg);
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < R; ++k) {
        if (fused[k]) {
            const int64_t rp = uni64(s_rp[brow[k]]);
            int32_t *cc = fo.ccol + rp;
            double *cv = fo.cval + rp;
            uint32_t before = 0;
#pragma unroll
            for (int j = 0; j < kFItems; ++j) {
                const uint32_t b = (out[k].exnz >> (8 * j)) & 0xffu;
                if (b & 0x80u) {
                    const uint32_t o = before + (b & 0x7fu);
                    __builtin_nontemporal_store(
                        out[k].col[j], reinterpret_cast<int32_t *>(reinterpret_cast<char *>(cc) + o * 4u));
                    __builtin_nontemporal_store(
                        out[k].val[j], reinterpret_cast<double *>(reinterpret_cast<char *>(cv) + o * 8u));
                }
                before += (out[k].tot >> (8 * j)) & 0xffu;
            }
        }
    }
    if (lane == 0 && flags) atomicOr(&ws.ctrl->flags, (unsigned long long)flags);
}

}  // namespace

void launch_pack_tiles(const int64_t *part, int64_t nblocks, int64_t *tcnt, int64_t *toff, int64_t *tmap,
                       int64_t *tmp, Control *ctrl, hipStream_t stream) {
    if (nblocks <= 0) return;
    const unsigned g = (unsigned)min((nblocks + 255) / 256, (int64_t)4096);
    hipLaunchKernelGGL(k_pack_counts, dim3(g), dim3(256), 0, stream, part, nblocks, tcnt);
    launch_scan(tcnt, toff, nblocks, tmp, &ctrl->packed_tiles, stream);
    hipLaunchKernelGGL(k_pack_map, dim3(g), dim3(256), 0, stream, tcnt, toff, nblocks, tmap);
}

void launch_fused_direct(const Operands &op, const Workspace &ws, const FusedOut &fo, bool small_b,
                         int rows_per_wave, hipStream_t stream) {
    if (op.arows <= 0 || fo.ntiles <= 0) return;
    const dim3 grid((unsigned)fo.ntiles), block(kFusedTileTpb);
    if (rows_per_wave == 1) {
        if (small_b) hipLaunchKernelGGL((k_fused<true, 1, true>), grid, block, 0, stream, op, ws, fo);
        else hipLaunchKernelGGL((k_fused<false, 1, true>), grid, block, 0, stream, op, ws, fo);
    } else {
        if (small_b) hipLaunchKernelGGL((k_fused<true, 2, true>), grid, block, 0, stream, op, ws, fo);
        else hipLaunchKernelGGL((k_fused<false, 2, true>), grid, block, 0, stream, op, ws, fo);
    }
}

int64_t fused_tiles(int32_t rows, int rows_per_wave) {
    const int64_t tile =
        rows_per_wave == kFusedDealRows ? (int64_t)kFusedDealTile : (int64_t)(kFusedTileTpb / 64) * rows_per_wave;
    return ((int64_t)rows + tile - 1) / tile;
}

void launch_fused(const Operands &op, const Workspace &ws, const FusedOut &fo, bool small_b, int rows_per_wave,
                  hipStream_t stream) {
    if (op.arows <= 0) return;
    const dim3 grid((unsigned)fo.ntiles), block(kFusedTpb);
    if (rows_per_wave == kFusedPackRows) {
        if (fo.ntiles <= 0) return;
        if (small_b) hipLaunchKernelGGL((k_fused_packed<true>), grid, block, 0, stream, op, ws, fo);
        else hipLaunchKernelGGL((k_fused_packed<false>), grid, block, 0, stream, op, ws, fo);
    } else if (rows_per_wave == kFusedDealRows) {
        const dim3 dblock(kFusedDealTpb);
        if (small_b) hipLaunchKernelGGL((k_fused_dealt<true>), grid, dblock, 0, stream, op, ws, fo);
        else hipLaunchKernelGGL((k_fused_dealt<false>), grid, dblock, 0, stream, op, ws, fo);
    } else {
        const dim3 tblock(kFusedTileTpb);
        if (rows_per_wave == 1) {
            if (small_b) hipLaunchKernelGGL((k_fused<true, 1>), grid, tblock, 0, stream, op, ws, fo);
            else hipLaunchKernelGGL((k_fused<false, 1>), grid, tblock, 0, stream, op, ws, fo);
        } else {
            if (small_b) hipLaunchKernelGGL((k_fused<true, 2>), grid, tblock, 0, stream, op, ws, fo);
            else hipLaunchKernelGGL((k_fused<false, 2>), grid, tblock, 0, stream, op, ws, fo);
        }
    }
}

}  // namespace spg
