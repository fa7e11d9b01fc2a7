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
// entries (R-MAT's first rows hold hundreds of thousands) is walked by all
// waves of its workgroup together, each taking a contiguous share of its
// entries, so no single wave carries it.
constexpr int64_t kAnHeavy = 16384;

// Entries [S, T) of a row tile (lane l holds the tile's row l, starting at
// entry a0; rows 0..last exist).  DIR32: B row k is [bdir32[k], bdir32[k+1])
// (B's own row_ptr, 32-bit copy).  Writes each entry's (B row start, length)
// and adds the lengths into s_ub[row] (LDS atomics: several waves may walk
// one tile).  s_mark is this wave's scratch.
template <bool DIR32>
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
            if (inv[h]) atomicAdd(&s_ub[row], (unsigned long long)lv[h]);
        }
    }
    return flags;
}

// Classification of a tile's rows from their summed bounds: ub, bins, bin
// lists and the wave's partial sums.
__device__ __forceinline__ void tile_finish(const Operands &op, const Workspace &ws, int64_t tile, int64_t a0,
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
        if (!m) continue;
        if (lane == 0) s_part[2 + bb] = (int64_t)__popcll(m);
        const bool listed = bb != kBinEmpty && ((ws.mode & kWsListSmall) || (bb != kBinW1 && bb != kBinW2));
        if (!listed) continue;
        int basei = 0;
        if (lane == 0) basei = atomicAdd(&ws.ctrl->bin_count[bb], (int)__popcll(m));
        basei = __shfl(basei, 0, 64);
        if (b == bb) {
            RowRec rec;
            rec.row = (int32_t)r;
            rec.na = (int32_t)min(a1 - a0, (int64_t)INT_MAX);
            rec.a0 = a0;
            ws.bin_recs[(size_t)bb * (size_t)ws.list_stride + (size_t)(basei + (int)__popcll(m & lanemask_lt()))] = rec;
        }
    }
}

template <bool DIR32>
__global__ __launch_bounds__(kAnalyzeTpb) void k_analyze(Operands op, Workspace ws) {
    constexpr int NW = kAnalyzeTpb / 64;
    __shared__ unsigned long long s_ub[NW][64];
    __shared__ int s_mark[NW][64];
    __shared__ int64_t s_part[NW][kAnalyzeFields];
    __shared__ int64_t s_span[NW][2];  // entry range of each wave's tile when heavy, else empty
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
    const int64_t E0 = __shfl(a0, 0, 64);
    const int64_t E1 = __shfl(a1, last, 64);
    s_ub[w][lane] = 0ull;
    unsigned flags = 0;
    if (has && op.nnz_a <= 0 && E1 > E0) flags |= kErrInternal;  // row_ptr inconsistent with nnz(A)
    const bool heavy = has && E1 - E0 > kAnHeavy;
    if (lane == 0) {
        s_span[w][0] = heavy ? E0 : 0;
        s_span[w][1] = heavy ? E1 : 0;
    }
    wave_sync();
    if (has && !heavy) flags |= tile_entries<DIR32>(op, ws, last, a0, E0, E1, s_ub[w], s_mark[w]);
    __syncthreads();
    for (int h = 0; h < NW; ++h) {  // heavy tiles of the workgroup, walked by every wave
        const int64_t hE0 = s_span[h][0], hE1 = s_span[h][1];
        if (hE1 <= hE0) continue;  // uniform across the workgroup
        const int64_t htile = (int64_t)blockIdx.x * NW + h;
        const int64_t hr = htile * 64 + lane;
        const int hlast = (int)min((int64_t)63, (int64_t)op.arows - 1 - htile * 64);
        const int64_t ha0 = hr < op.arows ? op.aptr[hr] : 0;
        const int64_t per = ((hE1 - hE0 + NW - 1) / NW + 63) / 64 * 64;
        const int64_t S = hE0 + (int64_t)w * per, T = min(S + per, hE1);
        if (S < T) flags |= tile_entries<DIR32>(op, ws, hlast, ha0, S, T, s_ub[h], s_mark[w]);
    }
    __syncthreads();
    if (has) tile_finish(op, ws, tile, a0, a1, s_ub[w], s_part[w]);
    unsigned long long fl = flags;
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) fl |= __shfl_xor(fl, d, 64);
    if (lane == 0 && fl) atomicOr(&ws.ctrl->flags, fl);
    __syncthreads();
    if (threadIdx.x < kAnalyzeFields) {
        int64_t v = 0;
#pragma unroll
        for (int i = 0; i < NW; ++i) v += s_part[i][threadIdx.x];
        ws.part[(size_t)blockIdx.x * kAnalyzeFields + threadIdx.x] = v;
    }
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
    if (ws.bdir32)
        hipLaunchKernelGGL(k_analyze<true>, dim3((unsigned)blocks), dim3(kAnalyzeTpb), 0, stream, op, ws);
    else
        hipLaunchKernelGGL(k_analyze<false>, dim3((unsigned)blocks), dim3(kAnalyzeTpb), 0, stream, op, ws);
    hipLaunchKernelGGL(k_analyze_reduce, dim3(1), dim3(kReduceTpb), 0, stream, ws.part, blocks, ws);
}

}  // namespace spg
