// Tile-binning helpers shared by the splat kernels (Born-Wolf tables and the
// separable Gaussian model): integer fills, the tile -> frame map and an
// in-LDS bitonic sort used to put each tile's emitter list in index order.
#pragma once

#include "common.hpp"
#include "scopyon_hip_types.h"

namespace smx {
// internal linkage: every translation unit gets its own copy of the kernels
namespace {

__global__ void fill_int_kernel(int* __restrict__ p, size_t n, int value) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

// tile_fov[g] = frame owning global tile g (-1 where no frame owns it).
__global__ void tile_fov_kernel(int nfov, const smx_fov_t* __restrict__ fovs,
                                int* __restrict__ tile_fov, int n_tiles) {
    int f = blockIdx.x;
    if (f >= nfov) return;
    const smx_fov_t F = fovs[f];
    int nt = F.tiles_r * F.tiles_c;
    for (int t = threadIdx.x; t < nt; t += blockDim.x) {
        int g = F.tile_off + t;
        if (g >= 0 && g < n_tiles) tile_fov[g] = f;
    }
}

// Per-tile counters are spaced `cs` ints apart (1, or COUNT_STRIDE = one
// 128-byte line).  Binning that makes one device atomic per entry serialises
// on shared lines when tiles hold many entries, so it pads them then; with
// few entries the padding only spreads the working set.
constexpr int COUNT_STRIDE = 32;
inline int count_stride(long long entries, int n_tiles, int dense_per_tile = 128) {
    return entries >= (long long)dense_per_tile * n_tiles ? COUNT_STRIDE : 1;
}
constexpr int BIN_WIN = 4096;
// Binning aggregated per block in LDS over a window of BIN_WIN tiles makes one
// device atomic per block and tile instead: when the window holds every tile
// those are the only atomics, and dense counters take them faster (256
// blocks x 4096 tiles, one atomic each: 12 us dense, 41 us padded on MI355X;
// 1,000,000 Gaussian emitters on 2048^2: binning + forward 0.214 -> 0.165 ms).
inline int agg_count_stride(long long entries, int n_tiles, int dense_per_tile = 128) {
    return n_tiles <= BIN_WIN ? 1 : count_stride(entries, n_tiles, dense_per_tile);
}

// Tile lists longer than a tile pass sorts in LDS (long_min entries) are
// noted while the counters are scanned: long_tiles[0] = how many,
// long_tiles[1 ..] = the tiles, in no particular order.  sort_long_lists_kernel
// then sorts each of them in place before the tile pass, which reads such a
// list in order from memory.
__device__ inline void note_long(int* long_tiles, int long_min, int g, int count) {
    if (long_tiles && count > long_min) {
        const int k = atomicAdd(long_tiles, 1);
        long_tiles[1 + k] = g;
    }
}

// Dense copy of the counters for the scan: dense[g] = counts[g * cs].
__global__ void compact_counts_kernel(const int* __restrict__ counts, int cs, int n,
                                      int* __restrict__ dense, int long_min,
                                      int* __restrict__ long_tiles) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) {
        const int c = counts[(size_t)g * cs];
        dense[g] = c;
        note_long(long_tiles, long_min, g, c);
    }
}

// Exclusive scan of the (strided) tile counters in one workgroup of
// SCAN_THREADS threads: begin[g] = sum of counts[h * cs] for h < g, g < n.
// The counters go in rows of SCAN_THREADS (thread t holds counter
// row * SCAN_THREADS + t: every load and store of a wave is coalesced),
// SCAN_ROWS rows at a time: the rows' wave scans are independent shuffle
// chains that overlap, their wave totals meet in LDS behind one barrier, and
// each row adds the totals of the waves before it and of the rows before it.
// Replaces compact + device-wide scan (three launches) up to SCAN_MAX
// counters.  (With a barrier per row, or a run of consecutive counters per
// thread, 16,385 counters took 13 us on MI355X against 2.3 us for an empty
// launch.)
constexpr int SCAN_THREADS = 1024;
constexpr int SCAN_ROWS = 8;
constexpr int SCAN_MAX = 65536;
__global__ void __launch_bounds__(SCAN_THREADS) scan_counts_kernel(const int* __restrict__ counts,
                                                                  int cs, int n,
                                                                  int* __restrict__ begin,
                                                                  int long_min,
                                                                  int* __restrict__ long_tiles) {
    constexpr int NW = SCAN_THREADS / 64;
    static_assert(NW == 16, "one DPP row of 16 lanes scans the wave totals");
    __shared__ int s_wave[SCAN_ROWS][NW];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    // the long-list count starts at zero here (a binning that skips the
    // counter initialisation relies on it)
    if (t == 0 && long_tiles) long_tiles[0] = 0;
    __syncthreads();
    int carry = 0;   // sum of the rows before
    for (int base = 0; base < n; base += SCAN_THREADS * SCAN_ROWS) {
        int c[SCAN_ROWS], incl[SCAN_ROWS];
#pragma unroll
        for (int r = 0; r < SCAN_ROWS; ++r) {
            const int g = base + r * SCAN_THREADS + t;
            c[r] = g < n ? counts[g * cs] : 0;   // g * cs < SCAN_MAX * COUNT_STRIDE
        }
#pragma unroll
        for (int r = 0; r < SCAN_ROWS; ++r) {
            const int g = base + r * SCAN_THREADS + t;
            if (g < n) note_long(long_tiles, long_min, g, c[r]);
            incl[r] = c[r];
        }
#pragma unroll
        for (int r = 0; r < SCAN_ROWS; ++r) incl[r] = wave_incl_scan(incl[r]);
        if (base > 0) __syncthreads();   // the previous rows' totals are read
        if (lane == 63) {
#pragma unroll
            for (int r = 0; r < SCAN_ROWS; ++r) s_wave[r][wave] = incl[r];
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < SCAN_ROWS; ++r) {
            // lanes 0..15 of every wave scan the 16 wave totals (one DPP row)
            int ws = s_wave[r][lane & (NW - 1)];
            ws += __builtin_amdgcn_update_dpp(0, ws, 0x111, 0xf, 0xf, false);   // row_shr:1
            ws += __builtin_amdgcn_update_dpp(0, ws, 0x112, 0xf, 0xf, false);   // row_shr:2
            ws += __builtin_amdgcn_update_dpp(0, ws, 0x114, 0xf, 0xf, false);   // row_shr:4
            ws += __builtin_amdgcn_update_dpp(0, ws, 0x118, 0xf, 0xf, false);   // row_shr:8
            const int before = wave > 0 ? __builtin_amdgcn_readlane(ws, wave - 1) : 0;
            const int total = __builtin_amdgcn_readlane(ws, NW - 1);
            const int g = base + r * SCAN_THREADS + t;
            if (g < n) begin[g] = carry + before + incl[r] - c[r];
            carry += total;
        }
    }
}

// One pass before binning: tile_count[g * cs] = 0 for g <= n_tiles and
// tile_fov[g] = the frame owning global tile g (-1 where none does).  The frames' tile ranges
// are staged in LDS (up to TILES_INIT_FOVS of them) and searched by bisection
// when their offsets ascend, linearly otherwise.
constexpr int TILES_INIT_FOVS = 2048;
struct TilesInitLds {
    int off[TILES_INIT_FOVS], end[TILES_INIT_FOVS];
    int unsorted;
};
// body of tiles_init_kernel for workgroup `block` (256 threads)
__device__ inline void tiles_init_block(TilesInitLds& L, int block, int nfov,
                                        const smx_fov_t* __restrict__ fovs, int n_tiles,
                                        int* __restrict__ tile_count, int cs,
                                        int* __restrict__ tile_fov,
                                        int* __restrict__ long_tiles) {
    const int staged = nfov < TILES_INIT_FOVS ? nfov : TILES_INIT_FOVS;
    if (threadIdx.x == 0) L.unsorted = 0;
    __syncthreads();
    for (int f = threadIdx.x; f < staged; f += blockDim.x) {
        L.off[f] = fovs[f].tile_off;
        L.end[f] = fovs[f].tile_off + fovs[f].tiles_r * fovs[f].tiles_c;
    }
    __syncthreads();
    for (int f = threadIdx.x + 1; f < staged; f += blockDim.x)
        if (L.off[f] < L.end[f - 1]) L.unsorted = 1;
    __syncthreads();
    const int g = block * blockDim.x + threadIdx.x;
    if (g > n_tiles) return;
    tile_count[(size_t)g * cs] = 0;
    if (g == n_tiles) {
        if (long_tiles) long_tiles[0] = 0;
        return;
    }
    int owner = -1;
    if (!L.unsorted && staged == nfov) {
        int lo = 0, hi = nfov;   // last frame with off <= g
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (L.off[mid] <= g) lo = mid;
            else hi = mid;
        }
        if (nfov > 0 && g >= L.off[lo] && g < L.end[lo]) owner = lo;
    } else {
        for (int f = 0; f < nfov; ++f) {
            const int off = f < staged ? L.off[f] : fovs[f].tile_off;
            const int end = f < staged ? L.end[f] : off + fovs[f].tiles_r * fovs[f].tiles_c;
            if (g >= off && g < end) {
                owner = f;
                break;
            }
        }
    }
    tile_fov[g] = owner;
}

__global__ void __launch_bounds__(256) tiles_init_kernel(int nfov,
                                                         const smx_fov_t* __restrict__ fovs,
                                                         int n_tiles, int* __restrict__ tile_count,
                                                         int cs, int* __restrict__ tile_fov,
                                                         int* __restrict__ long_tiles) {
    __shared__ TilesInitLds L;
    tiles_init_block(L, blockIdx.x, nfov, fovs, n_tiles, tile_count, cs, tile_fov, long_tiles);
}

// Ascending bitonic sort of n_pow2 ints in LDS by the whole workgroup.
__device__ inline void lds_bitonic_sort(int* a, int n_pow2) {
    for (int k = 2; k <= n_pow2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < n_pow2; i += blockDim.x) {
                int ixj = i ^ j;
                if (ixj > i) {
                    int x = a[i], y = a[ixj];
                    bool up = (i & k) == 0;
                    if ((x > y) == up) {
                        a[i] = y;
                        a[ixj] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// Ascending bitonic sort of up to 64 * NPL ints by one wave, in registers:
// element p = r * 64 + lane lives in v[r] (entries past `count` padded with
// INT_MAX).  Stage (k, j) pairs p with p ^ j -- across lanes for j < 64 (one
// shuffle), inside the lane for j >= 64 -- and the lower of the pair keeps
// the minimum when p & k == 0, the maximum otherwise.  The sorted list is
// written to out[0 .. 64 * NPL).
// Value of lane (lane xor j) for the sorts below: DPP lane moves for j < 16
// (quad_perm [1,0,3,2] and [2,3,0,1]; for 4 and 8 a quad or row reversal
// followed by the half-row mirror, since 7 - i = i ^ 7 inside 8 lanes and
// 15 - i = i ^ 15 inside 16), a shuffle (ds_bpermute) for 16 and 32.  j is a
// compile-time constant after unrolling; every lane of the wave is active.
__device__ inline int lane_xor(int v, int j) {
    if (j == 1) return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xf, 0xf, false);
    if (j == 2) return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xf, 0xf, false);
    if (j == 4)
        return __builtin_amdgcn_update_dpp(
                0, __builtin_amdgcn_update_dpp(0, v, 0x1B, 0xf, 0xf, false), 0x141, 0xf, 0xf, false);
    if (j == 8)
        return __builtin_amdgcn_update_dpp(
                0, __builtin_amdgcn_update_dpp(0, v, 0x140, 0xf, 0xf, false), 0x141, 0xf, 0xf, false);
    return __shfl_xor(v, j);
}

template <int NPL>
__device__ inline void wave_sort(const int* src, int count, int lane, int* out) {
    int v[NPL];
#pragma unroll
    for (int r = 0; r < NPL; ++r) {
        const int p = r * 64 + lane;
        v[r] = p < count ? src[p] : 0x7fffffff;
    }
#pragma unroll
    for (int k = 2; k <= 64 * NPL; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j >= 64) {
                const int m = j >> 6;
#pragma unroll
                for (int r = 0; r < NPL; ++r) {
                    if (r & m) continue;
                    const bool up = ((r * 64 + lane) & k) == 0;
                    const int lo = min(v[r], v[r | m]), hi = max(v[r], v[r | m]);
                    v[r] = up ? lo : hi;
                    v[r | m] = up ? hi : lo;
                }
            } else {
                const bool lower = (lane & j) == 0;
#pragma unroll
                for (int r = 0; r < NPL; ++r) {
                    // DPP moves when several registers interleave; one element
                    // per lane waits on their hazards (measured slower)
                    const int b = NPL >= 2 ? lane_xor(v[r], j) : __shfl_xor(v[r], j);
                    const bool up = ((r * 64 + lane) & k) == 0;
                    v[r] = (lower == up) ? min(v[r], b) : max(v[r], b);
                }
            }
        }
    }
#pragma unroll
    for (int r = 0; r < NPL; ++r) out[r * 64 + lane] = v[r];
}

// Sort of a tile list of 2 <= count <= 64 * NPL * (waves of the workgroup)
// distinct entries (< INT_MAX) by the whole workgroup: wave w sorts entries
// w RL .. w RL + RL - 1 (RL = 64 NPL) in registers into runs[w RL ..]
// (wave_sort), then an entry's place is its index in its own run plus, for
// every other run, how many of that run's entries are below it (a binary
// search; the INT_MAX padding never is).  runs[] holds (count rounded up to
// RL) ints, out[] count ints, not overlapping.  The same permutation as a rank
// sort, with (runs - 1) log2(2 RL) LDS reads per entry instead of count / 4.
template <int NPL>
__device__ inline void merge_rank_sort(const int* __restrict__ list, int count, int* runs,
                                       int* out) {
    constexpr int RL = 64 * NPL;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int nruns = (count + RL - 1) / RL;
    const int own = min(RL, count - w * RL);   // entries of this wave's run (<= 0: none)
    if (w < nruns) wave_sort<NPL>(list + w * RL, own, lane, runs + w * RL);
    __syncthreads();
    if (w < nruns) {
#pragma unroll
        for (int r = 0; r < NPL; ++r) {
            const int i = r * 64 + lane;
            if (i >= own) continue;
            const int v = runs[w * RL + i];
            int place = i;
            for (int o = 0; o < nruns; ++o) {
                if (o == w) continue;
                const int* b = runs + o * RL;
                int below = 0;   // entries of run o below v
#pragma unroll
                for (int step = RL; step > 0; step >>= 1) {
                    const int idx = below + step - 1;
                    if (idx < RL && b[idx] < v) below += step;
                }
                place += below;
            }
            out[place] = v;
        }
    }
    __syncthreads();
}

// Rank sort of a tile list of 2 <= count distinct entries by the whole
// workgroup: every entry's rank is the number of smaller entries (broadcast
// LDS reads, no dependent shuffle stages), so a list of up to one entry per
// thread sorts in one pass over it.  raw[] must hold count rounded up to 4
// and not overlap out[]; out[rank] = entry.
__device__ inline void rank_sort(const int* __restrict__ list, int count, int* raw, int* out) {
    const int c4 = (count + 3) & ~3;
    for (int k = threadIdx.x; k < c4; k += blockDim.x) raw[k] = k < count ? list[k] : 0x7fffffff;
    __syncthreads();
    const int4* raw4 = reinterpret_cast<const int4*>(raw);
    for (int p = threadIdx.x; p < count; p += blockDim.x) {
        const int v = raw[p];
        int r = 0;
        for (int q = 0; q < (c4 >> 2); ++q) {
            const int4 w = raw4[q];
            r += (w.x < v) + (w.y < v) + (w.z < v) + (w.w < v);
        }
        out[r] = v;
    }
    __syncthreads();
}

// Sort a tile list of `count` <= WAVE_SORT_MAX entries into out[] by one
// wave, in registers; out[] must hold the count rounded up to 128, 256, 512
// or 1024 (the tail is INT_MAX).
constexpr int WAVE_SORT_MAX = 1024;
__device__ inline void wave_sort_list(const int* src, int count, int lane, int* out) {
    if (count <= 128) wave_sort<2>(src, count, lane, out);
    else if (count <= 256) wave_sort<4>(src, count, lane, out);
    else if (count <= 512) wave_sort<8>(src, count, lane, out);
    else wave_sort<16>(src, count, lane, out);
}

// In-place sort of the noted long tile lists (see note_long), up to
// LONG_SORT_CAP entries each: every wave sorts 1024-entry runs in registers
// (wave_sort), then the runs are merged pairwise in LDS -- an entry's place in
// the merged run is its index in its own run plus its rank in the partner run
// (lower bound from the left run, upper bound from the right one, so the INT_MAX
// padding of the last run keeps distinct places).  A longer list stays in
// binning order and sets status bit 1.  Blocks take the noted tiles in a
// fixed stride; with none noted every block returns at once.
constexpr int LONG_SORT_THREADS = 512;
constexpr int LONG_SORT_CAP = 8192;
constexpr int LONG_SORT_BLOCKS = 256;
__global__ void __launch_bounds__(LONG_SORT_THREADS) sort_long_lists_kernel(
        const int* __restrict__ long_tiles, const int* __restrict__ tile_begin,
        int* __restrict__ list, long long capacity, int* __restrict__ status) {
    __shared__ int buf[2][LONG_SORT_CAP];
    const int nlong = long_tiles[0];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int k = blockIdx.x; k < nlong; k += gridDim.x) {
        const int t = long_tiles[1 + k];
        const int begin = tile_begin[t];
        int count = tile_begin[t + 1] - begin;
        if ((long long)begin + count > capacity) count = (int)max(0LL, capacity - begin);
        if (count > LONG_SORT_CAP) {
            if (threadIdx.x == 0) atomicOr(status, 1);
            continue;
        }
        int* L = list + begin;
        const int nch = (count + 1023) >> 10;
        for (int c = wave; c < nch; c += LONG_SORT_THREADS / 64)
            wave_sort<16>(L + c * 1024, min(1024, count - c * 1024), lane, buf[0] + c * 1024);
        __syncthreads();
        const int n2 = nch << 10;
        int src = 0;
        for (int run = 1024; run < n2; run <<= 1) {
            const int* a = buf[src];
            int* out = buf[src ^ 1];
            for (int i = threadIdx.x; i < n2; i += blockDim.x) {
                const int r = i / run, start = r * run;
                const int pstart = (r ^ 1) * run;
                const int plen = max(0, min(run, n2 - pstart));
                const int x = a[i];
                const int* p = a + pstart;
                int lo = 0, hi = plen;
                if ((r & 1) == 0) {
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if (p[mid] < x) lo = mid + 1;
                        else hi = mid;
                    }
                } else {
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if (p[mid] <= x) lo = mid + 1;
                        else hi = mid;
                    }
                }
                out[(r & ~1) * run + (i - start) + lo] = x;
            }
            __syncthreads();
            src ^= 1;
        }
        for (int i = threadIdx.x; i < count; i += blockDim.x) L[i] = buf[src][i];
        __syncthreads();
    }
}

// Binning with per-block aggregation: a block of BIN_THREADS emitters counts
// (and later ranks) its tile entries in LDS over a window of BIN_WIN tiles
// from `wbase`, then touches each window tile's global counter once, so the
// device-scope atomics per tile drop from one per entry to one per block.
// Entries outside the window use the global counter directly.
constexpr int BIN_THREADS = 1024;

__device__ inline void bin_hist_clear(int* s_hist) {
    for (int k = threadIdx.x; k < BIN_WIN; k += blockDim.x) s_hist[k] = 0;
}
// count one entry of `tile`
__device__ inline void bin_count(int* s_hist, int wbase, int tile, int* __restrict__ tile_count,
                                 int cs) {
    const unsigned d = (unsigned)(tile - wbase);
    if (d < (unsigned)BIN_WIN) atomicAdd(&s_hist[d], 1);
    else atomicAdd(&tile_count[(size_t)tile * cs], 1);
}
// after a barrier: add the block's window counts to the global counters
__device__ inline void bin_count_flush(const int* s_hist, int wbase, int* __restrict__ tile_count,
                                       int cs) {
    for (int k = threadIdx.x; k < BIN_WIN; k += blockDim.x) {
        const int c = s_hist[k];
        if (c) atomicAdd(&tile_count[(size_t)(wbase + k) * cs], c);
    }
}
// after a barrier: claim each window tile's slots from the counters, which
// run down to 0 -- the block's entries of tile wbase + k take [s_hist[k],
// s_hist[k] + count) on return
__device__ inline void bin_claim(int* s_hist, int wbase, int* __restrict__ tile_left, int cs) {
    for (int k = threadIdx.x; k < BIN_WIN; k += blockDim.x) {
        const int c = s_hist[k];
        if (c) s_hist[k] = atomicSub(&tile_left[(size_t)(wbase + k) * cs], c) - c;
    }
}

__device__ inline void lds_add(double* p, double v) {
    // ds_add_f64: one IEEE add in LDS, in program order per wave
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ inline void lds_add(float* p, float v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace
}  // namespace smx
