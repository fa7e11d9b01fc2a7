// analyze.hip -- per-row work bound, bins and bin row lists.
//
// P_row = sum over the row's A entries of nnz(B row) is the row's upper bound
// on distinct output columns and its share of scalarMultiplicationCount
// (src/common/spgemm_common.cpp:474-487).  One wave takes a tile of 64
// consecutive rows and streams the tile's A entries with all 64 lanes
// (coalesced, balanced whatever the row lengths), finds each entry's row with a
// prefix maximum over the row starts that fall in its 64-entry chunk, and sums
// lengths per row with LDS atomics.  Rows are appended to their bin's list with one
// atomic per bin per wave.  The (start, length) of the B row every A entry
// references is written out as a coalesced stream, so the later phases never
// gather B's row offsets again.
#include "device_common.hpp"

#include <algorithm>

namespace spg {
namespace {

using namespace dev;

__global__ void k_zero_control(Control *c) {
    int *p = reinterpret_cast<int *>(c);
    for (int i = threadIdx.x; i < (int)(sizeof(Control) / sizeof(int)); i += blockDim.x) p[i] = 0;
}

__global__ void k_zero_words(uint32_t *p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

constexpr int kAnalyzeTpb = 256;
static_assert(kAnalyzeTpb / 64 * 64 == kAnalyzeRows, "packed single-pass tiles take an analyze block's rows");
constexpr int kAnChunks = 2;  // 64-entry chunks in flight per wave (1 and 4 measured no faster)
// Heavy tiles are the critical path of products with power-law rows: they
// are cut into pieces of about kAnPiece entries that a separate kernel takes
// on demand, each piece walked by the 16 waves of a 1024-thread workgroup,
// each wave keeping more chunks in flight.
constexpr int kAnHeavyChunks = 4;
constexpr int kAnHeavyTpb = 1024;
constexpr int64_t kAnPiece = 32768;

// Longest row of A and of B (block maxima, one atomic per block each) and
// B's row offsets in 32 bits.
__global__ void k_rowmax(const int64_t *aptr, int32_t arows, const int64_t *bptr, int32_t brows, uint32_t *dir32,
                         Control *ctrl) {
    __shared__ unsigned long long s_m[2][4];
    unsigned long long ma = 0, mb = 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= (int64_t)brows; i += stride) {
        const int64_t v = bptr[i];
        dir32[i] = (uint32_t)v;
        if (i < brows) {
            const int64_t l = bptr[i + 1] - v;
            mb = max(mb, (unsigned long long)(l > 0 ? l : 0));
        }
    }
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (int64_t)arows; i += stride) {
        const int64_t l = aptr[i + 1] - aptr[i];
        ma = max(ma, (unsigned long long)(l > 0 ? l : 0));
    }
    for (int d = 32; d > 0; d >>= 1) {
        ma = max(ma, (unsigned long long)__shfl_xor((long long)ma, d, 64));
        mb = max(mb, (unsigned long long)__shfl_xor((long long)mb, d, 64));
    }
    const int w = threadIdx.x >> 6;
    if (lane_id() == 0) {
        s_m[0][w] = ma;
        s_m[1][w] = mb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < (int)(blockDim.x >> 6); ++k) {
            ma = max(ma, s_m[0][k]);
            mb = max(mb, s_m[1][k]);
        }
        if (ma) atomicMax(&ctrl->max_a_row, ma);
        if (mb) atomicMax(&ctrl->max_b_row, mb);
    }
}

__global__ void k_dir32(const int64_t *bptr, int64_t n, uint32_t *out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (uint32_t)bptr[i];
}

// Row tiles of 64 rows, one per wave.  A tile with more than kAnHeavy
// entries (R-MAT's first rows hold hundreds of thousands) is left to
// k_analyze_heavy.
constexpr int64_t kAnHeavy = 16384;

// Pieces of a heavy tile of E entries and r rows: at most one per row, so
// that the pieces of all heavy tiles fit a list of one entry per row of A.
__device__ __forceinline__ int64_t heavy_pieces(int64_t E, int rows) {
    const int64_t p = (E + kAnPiece - 1) / kAnPiece;
    return p < rows ? p : rows;
}

// Entries [S, T) of a row tile (lane l holds the tile's row l, starting at
// entry a0; rows 0..last exist).  DIR32: B row k is [bdir32[k], bdir32[k+1])
// (B's own row_ptr, 32-bit copy).  Writes each entry's (B row start, length)
// and adds the lengths into s_ub[row] (LDS atomics: several waves may walk
// one tile).  s_mark is this wave's scratch.
template <bool DIR32, int kAnChunks>
__device__ __forceinline__ unsigned tile_entries(const Operands &op, const Workspace &ws, int last, int64_t a0,
                                                 int64_t S, int64_t T, unsigned long long *s_ub, int *s_mark) {
    const int lane = lane_id();
    unsigned flags = 0;
    // kAnChunks chunks of 64 entries per step; every chunk's column loads and
    // B-row bound gathers are issued (unconditionally, at clamped indices)
    // before any chunk is reduced.
    for (int64_t base = S; op.nnz_a > 0 && base < T; base += 64 * kAnChunks) {
        int64_t ev[kAnChunks], sv[kAnChunks], lv[kAnChunks];
        bool inv[kAnChunks];
        int32_t kv[kAnChunks];
#pragma unroll
        for (int h = 0; h < kAnChunks; ++h) {
            ev[h] = base + 64 * h + lane;
            inv[h] = ev[h] < T;
            if (inv[h] && ev[h] >= op.nnz_a) {  // row_ptr inconsistent with nnz(A)
                flags |= kErrInternal;
                inv[h] = false;
            }
            const int64_t ec = max((int64_t)0, min(ev[h], min(T, op.nnz_a) - 1));
            kv[h] = __builtin_nontemporal_load(&op.acol[ec]);
        }
        int64_t bs[kAnChunks], be[kAnChunks];
#pragma unroll
        for (int h = 0; h < kAnChunks; ++h) {
            const bool ok = kv[h] >= 0 && kv[h] < op.ndir;
            const int32_t kk = ok ? kv[h] : 0;
            if (DIR32) {
                bs[h] = (int64_t)ws.bdir32[kk];
                be[h] = (int64_t)ws.bdir32[kk + 1];
            } else {
                bs[h] = op.ndir > 0 ? op.bstart[kk] : -1;
                be[h] = op.ndir > 0 ? op.bend[kk] : -1;
            }
            if (inv[h] && !ok) flags |= kErrRange;
            sv[h] = ok ? bs[h] : -1;
        }
#pragma unroll
        for (int h = 0; h < kAnChunks; ++h) {
            int64_t s0 = 0, len = 0;
            if (inv[h] && kv[h] >= 0 && kv[h] < op.ndir) {
                if (sv[h] < 0) {
                    flags |= kErrMissing;
                } else {
                    s0 = sv[h];
                    len = be[h] > s0 ? be[h] - s0 : 0;
                    if (len > INT_MAX) {
                        flags |= kErrLongRow;
                        len = 0;
                    }
                }
            }
            lv[h] = len;
            if (inv[h]) {
                if (ws.mode & kWsNarrow)
                    __builtin_nontemporal_store((uint32_t)s0, reinterpret_cast<uint32_t *>(ws.ebst) + ev[h]);
                else
                    __builtin_nontemporal_store(s0, &ws.ebst[ev[h]]);
                __builtin_nontemporal_store((int32_t)len, &ws.elen[ev[h]]);
            }
        }
#pragma unroll
        for (int h = 0; h < kAnChunks; ++h) {
            // Row (lane) holding entry e = base + 64h + lane: the largest row
            // l <= last whose first entry a0_l <= e.  Rows that start inside
            // the chunk mark their start position; a prefix maximum over the
            // positions, seeded with the last row starting at or before the
            // chunk, gives every entry's row.
            const int64_t cb = base + 64 * h;
            s_mark[lane] = -1;
            wave_sync();
            if (lane <= last && a0 > cb && a0 < cb + 64) atomicMax(&s_mark[a0 - cb], lane);
            const int seed = wave_max_dpp(lane <= last && a0 <= cb ? lane : -1);
            wave_sync();
            const int row = max(seed, wave_incl_max_dpp(s_mark[lane]));
            // The lengths of a row's lanes (consecutive) are summed in the
            // wave first: one LDS atomic per row of the chunk, not one per
            // entry (64 on one address inside a long row).  The inclusive
            // sum never decreases, so a run's base (the sum before its first
            // lane) is the prefix maximum of the bases marked at run starts;
            // every move is a DPP lane shift.
            const int64_t v = inv[h] ? lv[h] : 0;
            if (__ballot(v >= ((int64_t)1 << 24)) == 0) {
                const int incl = wave_incl_sum_dpp((int)v);
                const int prow = dpp_i<0x138, 0xf>(INT_MIN, row);  // lane - 1 (wave_shr 1); lane 0: none
                const int nrow = dpp_i<0x130, 0xf>(INT_MIN, row);  // lane + 1 (wave_shl 1); lane 63: none
                const int base = wave_incl_max_dpp(prow != row ? incl - (int)v : INT_MIN);
                const int sum = incl - base;
                if ((lane == 63 || nrow != row) && sum) atomicAdd(&s_ub[row], (unsigned long long)sum);
            } else if (inv[h]) {
                atomicAdd(&s_ub[row], (unsigned long long)lv[h]);
            }
        }
    }
    return flags;
}

// Classification of a tile's rows from their summed bounds: ub, bins and the
// wave's partial sums.  Returns the lane's bin (-1 past the last row).
__device__ __forceinline__ int tile_classify(const Operands &op, const Workspace &ws, int64_t tile, int64_t a0,
                                             int64_t a1, const unsigned long long *s_ub, int64_t *s_part) {
    const int lane = lane_id();
    const int64_t r = tile * 64 + lane;
    const bool valid = r < op.arows;
    const int64_t ub = (int64_t)s_ub[lane];
    const int b = valid ? classify(ub, a1 - a0) : -1;
    if (valid) {
        ws.ub[r] = ub;
        if (b == kBinEmpty) {
            ws.snnz[r] = 0;
            ws.annz[r] = 0;
        }
    }
    const int64_t tot = wave_sum(valid ? ub : (int64_t)0);
    const int64_t small = wave_sum((b == kBinW1 || b == kBinW2) ? ub : (int64_t)0);
    if (lane == 0) {
        s_part[0] = tot;
        s_part[1] = small;
    }
#pragma unroll
    for (int bb = 0; bb < kNumBins; ++bb) {
        const unsigned long long m = __ballot(b == bb);
        if (m && lane == 0) s_part[2 + bb] = (int64_t)__popcll(m);
    }
    return b;
}

// Whether bin bb's rows are listed (the single pass takes bins W1 and W2 by
// row range unless the lists are asked for).
__device__ __forceinline__ bool bin_listed(const Workspace &ws, int bb) {
    return bb != kBinEmpty && ((ws.mode & kWsListSmall) || (bb != kBinW1 && bb != kBinW2));
}

// The tile's rows of listed bins, in lane order from list position base[bb].
__device__ __forceinline__ void tile_list(const Workspace &ws, int64_t tile, int64_t a0, int64_t a1, int b,
                                          const int *base) {
    const int64_t r = tile * 64 + lane_id();
#pragma unroll
    for (int bb = 0; bb < kNumBins; ++bb) {
        const unsigned long long m = __ballot(b == bb);
        if (!m || !bin_listed(ws, bb)) continue;
        if (b == bb) {
            RowRec rec;
            rec.row = (int32_t)r;
            rec.na = (int32_t)min(a1 - a0, (int64_t)INT_MAX);
            rec.a0 = a0;
            ws.bin_recs[(size_t)bb * (size_t)ws.list_stride + (size_t)(base[bb] + (int)__popcll(m & lanemask_lt()))] =
                rec;
        }
    }
}

// A tile classified and listed by one wave alone: one list reservation per
// bin that holds rows.
__device__ __forceinline__ void tile_finish(const Operands &op, const Workspace &ws, int64_t tile, int64_t a0,
                                            int64_t a1, const unsigned long long *s_ub, int64_t *s_part) {
    const int b = tile_classify(op, ws, tile, a0, a1, s_ub, s_part);
    int base[kNumBins];
#pragma unroll
    for (int bb = 0; bb < kNumBins; ++bb) {
        const unsigned long long m = __ballot(b == bb);
        base[bb] = 0;
        if (!m || !bin_listed(ws, bb)) continue;
        int bi = 0;
        if (lane_id() == 0) bi = atomicAdd(&ws.ctrl->bin_count[bb], (int)__popcll(m));
        base[bb] = __builtin_amdgcn_readlane(bi, 0);
    }
    tile_list(ws, tile, a0, a1, b, base);
}

template <bool DIR32>
__global__ __launch_bounds__(kAnalyzeTpb) void k_analyze(Operands op, Workspace ws) {
    constexpr int NW = kAnalyzeTpb / 64;
    __shared__ unsigned long long s_ub[NW][64];
    __shared__ int s_mark[NW][64];
    __shared__ int64_t s_part[NW][kAnalyzeFields];
    __shared__ int s_lcnt[NW][kNumBins];  // listed rows per tile and bin, then the tile's first list position
    const int w = threadIdx.x >> 6;
    const int lane = lane_id();
    const int64_t tile = (int64_t)blockIdx.x * NW + w;
    const bool has = tile * 64 < op.arows;  // uniform per wave
    if (lane < kAnalyzeFields) s_part[w][lane] = 0;
    const int64_t r = tile * 64 + lane;
    const bool valid = r < op.arows;
    const int last = has ? (int)min((int64_t)63, (int64_t)op.arows - 1 - tile * 64) : 0;
    const int64_t a0 = valid ? op.aptr[r] : 0;
    const int64_t a1 = valid ? op.aptr[r + 1] : 0;
    const int64_t E0 = lane_value64(a0, 0);
    const int64_t E1 = lane_value64(a1, last);
    s_ub[w][lane] = 0ull;
    unsigned flags = 0;
    if (has && op.nnz_a <= 0 && E1 > E0) flags |= kErrInternal;  // row_ptr inconsistent with nnz(A)
    // Heavy tiles (their rows and their share of the block's partial sums)
    // are k_analyze_heavy's: the tile's rows start from 0 in ub, and its
    // pieces are listed in the storage of the empty bin's list (which lists
    // no rows; k_analyze_reduce reuses it after).  Record of piece k: row =
    // tile, na = k, a0 = index of piece 0 (piece 0: its pieces done, from 0).
    const bool heavy = has && E1 - E0 > kAnHeavy;
    if (heavy) {
        if (valid) ws.ub[r] = 0;
        const int np = (int)heavy_pieces(E1 - E0, last + 1);
        int base = 0;
        if (lane == 0) base = atomicAdd(&ws.ctrl->heavy_units, np);
        base = __builtin_amdgcn_readlane(base, 0);
        RowRec *hl = ws.bin_recs + (size_t)kBinEmpty * (size_t)ws.list_stride;
        for (int k = lane; k < np; k += 64) {
            RowRec rec;
            rec.row = (int32_t)tile;
            rec.na = k;
            rec.a0 = k == 0 ? 0 : base;
            hl[base + k] = rec;
        }
    }
    wave_sync();
    int b = -1;
    if (has && !heavy) {
        flags |= tile_entries<DIR32, kAnChunks>(op, ws, last, a0, E0, E1, s_ub[w], s_mark[w]);
        wave_sync();
        b = tile_classify(op, ws, tile, a0, a1, s_ub[w], s_part[w]);
    }
    // List positions are reserved once per bin for the block's four tiles
    // (one device-scope atomic on a bin's counter per block, not per wave).
#pragma unroll
    for (int bb = 0; bb < kNumBins; ++bb) {
        const unsigned long long m = __ballot(b == bb);
        if (lane == 0) s_lcnt[w][bb] = bin_listed(ws, bb) ? (int)__popcll(m) : 0;
    }
    __syncthreads();
    if (threadIdx.x < kNumBins) {
        int run = 0;
#pragma unroll
        for (int i = 0; i < NW; ++i) run += s_lcnt[i][threadIdx.x];
        int base = run ? atomicAdd(&ws.ctrl->bin_count[threadIdx.x], run) : 0;
#pragma unroll
        for (int i = 0; i < NW; ++i) {
            const int c = s_lcnt[i][threadIdx.x];
            s_lcnt[i][threadIdx.x] = base;
            base += c;
        }
    }
    __syncthreads();
    if (has && !heavy) tile_list(ws, tile, a0, a1, b, s_lcnt[w]);
    const unsigned fl = wave_or(flags);
    if (lane == 0 && fl) atomicOr(&ws.ctrl->flags, (unsigned long long)fl);
    if (threadIdx.x < kAnalyzeFields) {
        int64_t v = 0;
#pragma unroll
        for (int i = 0; i < NW; ++i) v += s_part[i][threadIdx.x];
        ws.part[(size_t)blockIdx.x * kAnalyzeFields + threadIdx.x] = v;
    }
}

// Pieces of heavy tiles, taken on demand.  The 16 waves walk contiguous
// shares of a piece's entries into LDS row sums, which go to ub with global
// atomics; the workgroup that completes a tile's last piece classifies the
// tile's rows and adds the tile's partial sums to its analyze block's (whose
// other tiles k_analyze has stored).
template <bool DIR32>
__global__ __launch_bounds__(kAnHeavyTpb) void k_analyze_heavy(Operands op, Workspace ws) {
    constexpr int NW = kAnHeavyTpb / 64;
    constexpr int TPW = kAnalyzeTpb / 64;  // tiles per analyze block
    __shared__ unsigned long long s_ub[64];
    __shared__ int s_mark[NW][64];
    __shared__ int64_t s_part[kAnalyzeFields];
    __shared__ int s_next, s_last;
    const int w = threadIdx.x >> 6;
    const int lane = lane_id();
    RowRec *hl = ws.bin_recs + (size_t)kBinEmpty * (size_t)ws.list_stride;
    const int units = ws.ctrl->heavy_units;
    unsigned flags = 0;
    for (int u; (u = next_large_row(&ws.ctrl->heavy_next, &s_next)) < units;) {
        const RowRec rec = hl[u];
        const int64_t tile = rec.row;
        const int first = rec.na == 0 ? u : (int)rec.a0;
        const int64_t r0 = tile * 64;
        const int last = (int)min((int64_t)63, (int64_t)op.arows - 1 - r0);
        const int64_t E0 = op.aptr[r0], E1 = op.aptr[r0 + last + 1];
        const int64_t np = heavy_pieces(E1 - E0, last + 1);
        const int64_t pp = ((E1 - E0 + np - 1) / np + 63) / 64 * 64;
        const int64_t P0 = min(E0 + (int64_t)rec.na * pp, E1), P1 = min(P0 + pp, E1);
        const bool valid = lane <= last;
        const int64_t a0 = valid ? op.aptr[r0 + lane] : 0;
        const int64_t a1 = valid ? op.aptr[r0 + lane + 1] : 0;
        if (threadIdx.x < 64) s_ub[threadIdx.x] = 0ull;
        __syncthreads();
        const int64_t per = ((P1 - P0 + NW - 1) / NW + 63) / 64 * 64;
        const int64_t S = min(P0 + (int64_t)w * per, P1), T = min(S + per, P1);
        if (S < T) flags |= tile_entries<DIR32, kAnHeavyChunks>(op, ws, last, a0, S, T, s_ub, s_mark[w]);
        __syncthreads();
        if (w == 0 && valid && s_ub[lane])
            atomicAdd(reinterpret_cast<unsigned long long *>(&ws.ub[r0 + lane]), s_ub[lane]);
        __threadfence();
        __syncthreads();
        if (threadIdx.x == 0)
            s_last = atomicAdd(reinterpret_cast<unsigned long long *>(&hl[first].a0), 1ull) ==
                     (unsigned long long)(np - 1);
        __syncthreads();
        if (s_last && w == 0) {
            __threadfence();
            s_ub[lane] = valid ? __hip_atomic_load(reinterpret_cast<const unsigned long long *>(&ws.ub[r0 + lane]),
                                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                               : 0ull;
            if (lane < kAnalyzeFields) s_part[lane] = 0;
            wave_sync();
            tile_finish(op, ws, tile, a0, a1, s_ub, s_part);
            wave_sync();
            if (lane < kAnalyzeFields && s_part[lane])
                atomicAdd(reinterpret_cast<unsigned long long *>(&ws.part[(size_t)(tile / TPW) * kAnalyzeFields + lane]),
                          (unsigned long long)s_part[lane]);
        }
        __syncthreads();  // piece state is reset for the next piece
    }
    const unsigned fl = wave_or(flags);
    if (lane == 0 && fl) atomicOr(&ws.ctrl->flags, (unsigned long long)fl);
}

// Large-row work classes: two per power of two of P_row (its leading bit
// and the next one), so rows of one class differ by less than 1.5x.
constexpr int kLargeClasses = 128;
__device__ __forceinline__ int large_class(int64_t ub) {
    const int lg = 63 - __clzll((long long)max(ub, (int64_t)2));
    return 2 * lg + (int)((ub >> (lg - 1)) & 1);
}

// Every field of every block's partials in one pass: thread i sums blocks
// i, i + 1024, ... for all fields at once.  Then the large rows are listed
// again in descending work class (heaviest first) in the storage of the
// empty bin, which lists no rows: the large-row kernels take rows in list
// order, so the heaviest start first and the lightest fill the tail.  The
// order inside a class is arbitrary; no result depends on it.
constexpr int kReduceTpb = 1024;
__global__ __launch_bounds__(kReduceTpb) void k_analyze_reduce(const int64_t *part, int64_t nblocks, Workspace ws) {
    Control *c = ws.ctrl;
    __shared__ int64_t s_f[kReduceTpb / 64][kAnalyzeFields];
    int64_t v[kAnalyzeFields] = {};
    for (int64_t i = threadIdx.x; i < nblocks; i += kReduceTpb) {
#pragma unroll
        for (int f = 0; f < kAnalyzeFields; ++f) v[f] += part[(size_t)i * kAnalyzeFields + f];
    }
#pragma unroll
    for (int f = 0; f < kAnalyzeFields; ++f) v[f] = wave_sum(v[f]);
    if (lane_id() == 0) {
#pragma unroll
        for (int f = 0; f < kAnalyzeFields; ++f) s_f[threadIdx.x >> 6][f] = v[f];
    }
    __syncthreads();
    const int f = threadIdx.x;
    if (f < kAnalyzeFields) {
        int64_t t = 0;
#pragma unroll
        for (int w = 0; w < kReduceTpb / 64; ++w) t += s_f[w][f];
        if (f == 0) c->products = (unsigned long long)t;
        else if (f == 1) c->ub_small = (unsigned long long)t;
        else c->bin_rows[f - 2] = t;
    }
    const int n = c->bin_count[kBinLarge];
    if (n <= 0) return;
    __shared__ int s_cls[kLargeClasses];
    const RowRec *in = ws.bin_recs + (size_t)kBinLarge * (size_t)ws.list_stride;
    RowRec *out = ws.bin_recs + (size_t)kBinEmpty * (size_t)ws.list_stride;
    if (threadIdx.x < kLargeClasses) s_cls[threadIdx.x] = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += kReduceTpb) atomicAdd(&s_cls[large_class(ws.ub[in[i].row])], 1);
    __syncthreads();
    if (threadIdx.x == 0) {  // class starts, heaviest class first
        int run = 0;
        for (int k = kLargeClasses - 1; k >= 0; --k) {
            const int cnt = s_cls[k];
            s_cls[k] = run;
            run += cnt;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += kReduceTpb) {
        const RowRec r = in[i];
        out[atomicAdd(&s_cls[large_class(ws.ub[r.row])], 1)] = r;
    }
}

}  // namespace

void launch_zero_bytes(void *p, size_t bytes, hipStream_t stream) {
    const size_t n = bytes / 4;
    if (!n) return;
    size_t blocks = (n + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    hipLaunchKernelGGL(k_zero_words, dim3((unsigned)blocks), dim3(256), 0, stream, static_cast<uint32_t *>(p), n);
}

void launch_zero_control(Control *ctrl, hipStream_t stream) {
    hipLaunchKernelGGL(k_zero_control, dim3(1), dim3(64), 0, stream, ctrl);
}

void launch_dir32(const int64_t *bptr, int64_t n, uint32_t *out, hipStream_t stream) {
    if (n <= 0) return;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_dir32, dim3((unsigned)blocks), dim3(256), 0, stream, bptr, n, out);
}

void launch_rowmax(const int64_t *aptr, int32_t arows, const int64_t *bptr, int32_t brows, uint32_t *dir32,
                   Control *ctrl, hipStream_t stream) {
    const int64_t n = std::max<int64_t>((int64_t)arows, (int64_t)brows + 1);
    int64_t blocks = (n + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_rowmax, dim3((unsigned)blocks), dim3(256), 0, stream, aptr, arows, bptr, brows, dir32, ctrl);
}

int64_t analyze_blocks(int32_t rows) {
    const int64_t tiles = ((int64_t)rows + 63) / 64;
    return (tiles + kAnalyzeTpb / 64 - 1) / (kAnalyzeTpb / 64);
}

void launch_analyze(const Operands &op, const Workspace &ws, hipStream_t stream) {
    if (op.arows <= 0) return;
    const int64_t blocks = analyze_blocks(op.arows);
    const unsigned gh = 256;  // workgroups taking heavy pieces (they exit at once when there are none)
    if (ws.bdir32) {
        hipLaunchKernelGGL(k_analyze<true>, dim3((unsigned)blocks), dim3(kAnalyzeTpb), 0, stream, op, ws);
        hipLaunchKernelGGL(k_analyze_heavy<true>, dim3(gh), dim3(kAnHeavyTpb), 0, stream, op, ws);
    } else {
        hipLaunchKernelGGL(k_analyze<false>, dim3((unsigned)blocks), dim3(kAnalyzeTpb), 0, stream, op, ws);
        hipLaunchKernelGGL(k_analyze_heavy<false>, dim3(gh), dim3(kAnHeavyTpb), 0, stream, op, ws);
    }
    hipLaunchKernelGGL(k_analyze_reduce, dim3(1), dim3(kReduceTpb), 0, stream, ws.part, blocks, ws);
}

}  // namespace spg
