// Separable Gaussian image model: forward image, Poisson negative
// log-likelihood with its image gradient, and the hand-written backward with
// respect to emitter positions, depths, intensities and the PSF width.
//
// Reference: the Gaussian radial profile exp(-r^2 / (2 w^2)) / (2 pi w^2) of
// src/scopyon/_epifm.py:128-134 (depth independent), integrated over each
// pixel; the TIRF amplitude exp(-|depth| / penetration_depth) of
// _epifm.py:1289.  Units here are pixels (x, y, sigma) and micrometres (z).
//
// Forward: emitters are binned to 32 x 32 tiles and each tile's list is
// sorted by emitter index (deterministic, no float atomics in global memory).
// The matrix-core form writes a tile as the contraction of per-emitter row
// and column factors (sep_gemm_kernel); the LDS form adds each emitter's
// products into LDS accumulators (gauss_forward_kernel).  Window factors come
// from a float64-built table over sub-pixel phases (gauss_table_kernel).
//
// Backward: every gradient is a separable sum over the emitter's window,
//   sum_ij g_ij u_i v_j  (g = dL/dmu, u, v in {G, dG/dx, dG/dsigma}),
// reduced by a 16-lane group per emitter: each lane owns one window column,
// walks the rows and the group combines five dot products with shuffles.
// The standalone form reads g from the image; the binned form reuses the
// forward's tile lists and reads each tile's window of g from LDS.
#include "common.hpp"
#include "scopyon_hip_types.h"
#include "tiles.hpp"

#include <hipcub/hipcub.hpp>
#include <algorithm>
#include <type_traits>

namespace {

using smx::align256;
using smx::fill_int_kernel;
using smx::lds_add;
using smx::lds_bitonic_sort;
using smx::wave_sort_list;
using smx::WAVE_SORT_MAX;
using smx::tile_fov_kernel;

constexpr int TS = 32;           // tile side
constexpr int QS = 16;           // quadrant side (one wave)
constexpr int FCHUNK = 64;       // emitters staged per round
constexpr int PSORT_CAP = 2048;  // tile list sorted in LDS
constexpr int WMAX = 15;         // largest window (radius 7)
constexpr int NSLOT = 8;         // 64-pixel parts in flight per wave
constexpr float INV_SQRT2 = 0.70710678118654752f;

// Pixel-integrated unit Gaussian along one axis: (erf(t1) - erf(t0)) / 2 with
// t = (edge - x) / (sqrt(2) sigma) for the pixel's two edges t0 < t1.  A pixel
// on one side of x takes the complementary form from e0 = erfc(|t0|) and
// e1 = erfc(|t1|), which keeps the tails accurate (no difference of two
// numbers close to +-1); the pixel containing x takes erf itself.  The edge
// values are shared with the neighbouring pixels by the callers.
__device__ inline float g_pixel_edges(float t0, float t1, float e0, float e1) {
    if (t0 >= 0.0f) return 0.5f * (e0 - e1);
    if (t1 <= 0.0f) return 0.5f * (e1 - e0);
    return 0.5f * (erff(t1) - erff(t0));
}

__device__ inline bool finite_coord(float v) { return v > -1.0e6f && v < 1.0e9f; }

// Window factors by sub-pixel phase.  For an emitter at x = floor(x) + d,
// window pixel b (of 2R + 1) spans [b - R - d, b - R + 1 - d) relative to x,
// so its factors depend on (d, b) only: G (the pixel-integrated unit
// Gaussian), dG/dx and dG/dsigma.  They are evaluated in float64 once per
// sigma at GQ + 1 phases d = q / GQ (gauss_table_kernel) and read with linear
// interpolation in d (error ~1e-8 of the peak factor at GQ = 1024 and sigma
// >= 0.5 px: the second derivative in d is bounded by 0.24 / sigma^2).  Entry
// (q, b) of each quantity holds the pair (value at q, value at q + 1 minus
// it), so one 8-byte load and one FMA give an interpolated factor; the three
// quantities are stored one after the other (G first: the forward reads only
// that 128 KB block).
constexpr int GQ = 1024;
constexpr int GB = 16;                   // factor slots per phase (window <= 15)
constexpr int TABLE_ENTRIES = GQ * GB;   // per quantity
constexpr int TABLE_PAIRS = 3 * TABLE_ENTRIES;

// G, dG/dx and dG/dsigma of window pixel b at phase d (float64)
__device__ inline void gauss_factors_f64(double sigma, int R, double d, int b, double out[3]) {
    const double e0 = (double)(b - R) - d, e1 = e0 + 1.0;
    const double s = 0.70710678118654752 / sigma;
    double g;
    if (e0 >= 0.0) g = 0.5 * (erfc(e0 * s) - erfc(e1 * s));
    else if (e1 <= 0.0) g = 0.5 * (erfc(-e1 * s) - erfc(-e0 * s));
    else g = 0.5 * (erf(e1 * s) - erf(e0 * s));
    const double c = 0.39894228040143268 / sigma;
    const double p0 = c * exp(-0.5 * (e0 / sigma) * (e0 / sigma));
    const double p1 = c * exp(-0.5 * (e1 / sigma) * (e1 / sigma));
    out[0] = g;
    out[1] = p0 - p1;
    out[2] = -(e1 * p1 - e0 * p0) / sigma;
}

// one table entry (q, b) of all three quantities; idx < TABLE_ENTRIES
__device__ inline void gauss_table_entry(const smx_gauss_params_t& P, int idx,
                                         float2* __restrict__ tab) {
    const int q = idx / GB, b = idx % GB;
    const int R = P.radius, W = 2 * R + 1;
    double v0[3] = {0.0, 0.0, 0.0}, v1[3] = {0.0, 0.0, 0.0};
    if (b < W) {
        gauss_factors_f64((double)P.sigma, R, (double)q / GQ, b, v0);
        gauss_factors_f64((double)P.sigma, R, (double)(q + 1) / GQ, b, v1);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float a = (float)v0[k];
        tab[k * TABLE_ENTRIES + idx] = make_float2(a, (float)(v1[k] - (double)a));
    }
}

__global__ void gauss_table_kernel(smx_gauss_params_t P, float2* __restrict__ tab) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < TABLE_ENTRIES) gauss_table_entry(P, idx, tab);
}

__device__ inline int table_slot(float d, float& f) {
    const float t = d * (float)GQ;
    const int q = min((int)t, GQ - 1);
    f = t - (float)q;
    return q * GB;
}

// G of window pixel b at sub-pixel offset d in [0, 1)
__device__ inline float table_g(const float2* __restrict__ tab, float d, int b) {
    float f;
    const float2 a = tab[table_slot(d, f) + b];
    return fmaf(f, a.y, a.x);
}

// (G, dG/dx, dG/dsigma) of window pixel b at sub-pixel offset d in [0, 1)
__device__ inline float4 table_factor(const float2* __restrict__ tab, float d, int b) {
    float f;
    const int o = table_slot(d, f) + b;
    const float2 g = tab[o], dx = tab[TABLE_ENTRIES + o], ds = tab[2 * TABLE_ENTRIES + o];
    return make_float4(fmaf(f, g.y, g.x), fmaf(f, dx.y, dx.x), fmaf(f, ds.y, ds.x), 0.0f);
}

struct TileRange { int r0, r1, c0, c1; };

// Window rows [floor(x) - R, floor(x) + R] and columns likewise, clipped to
// the frame; false when the window misses the frame.
__device__ inline bool window_tiles(const smx_fov_t& F, float x, float y, int R, TileRange& t) {
    if (!finite_coord(x) || !finite_coord(y)) return false;
    int lu = (int)floorf(x) - R, lv = (int)floorf(y) - R;
    int r0 = max(lu, 0), r1 = min(lu + 2 * R, F.nw - 1);
    int c0 = max(lv, 0), c1 = min(lv + 2 * R, F.nh - 1);
    if (r1 < r0 || c1 < c0) return false;
    t.r0 = r0 / TS; t.r1 = r1 / TS; t.c0 = c0 / TS; t.c1 = c1 / TS;
    return true;
}

__device__ inline int frame_of(const int* __restrict__ frame_id, int i) { return frame_id ? frame_id[i] : 0; }

// The tile that owns an emitter in the binned backward: the tile of its
// window centre pixel clamped into the frame.  The binning lists the emitter
// in that tile (the clamped centre lies inside the window's clipped range),
// so every listed emitter has exactly one owner; -1 when the window misses
// the frame or the tile lies outside the launch.
__device__ inline int owner_tile(const smx_fov_t& F, float x, float y, int R, int n_tiles) {
    TileRange t;
    if (!window_tiles(F, x, y, R, t)) return -1;
    const int cu = min(max((int)floorf(x), 0), F.nw - 1);
    const int cv = min(max((int)floorf(y), 0), F.nh - 1);
    const int tile = F.tile_off + (cu / TS) * F.tiles_c + cv / TS;
    return tile >= 0 && tile < n_tiles ? tile : -1;
}

// What a binning was built for (written by gauss_init_kernel, checked by the
// binned backward).
struct GaussHeader {
    float sigma;
    int radius;
    int n;
    int n_tiles;
};

// Tile counters and the tile -> frame map (blocks below tile_blocks), the
// phase table and the header (the blocks after them), in one launch.
__global__ void __launch_bounds__(256) gauss_init_kernel(int nfov,
                                                         const smx_fov_t* __restrict__ fovs,
                                                         int n_tiles, int* __restrict__ tile_count,
                                                         int cs, int* __restrict__ tile_fov,
                                                         int tile_blocks, smx_gauss_params_t P,
                                                         int n, float2* __restrict__ tab,
                                                         GaussHeader* __restrict__ hdr,
                                                         int* __restrict__ long_tiles) {
    __shared__ smx::TilesInitLds L;
    if ((int)blockIdx.x < tile_blocks) {
        smx::tiles_init_block(L, blockIdx.x, nfov, fovs, n_tiles, tile_count, cs, tile_fov,
                              long_tiles);
        return;
    }
    const int idx = ((int)blockIdx.x - tile_blocks) * blockDim.x + threadIdx.x;
    if (idx == 0) {
        GaussHeader h;
        h.sigma = P.sigma;
        h.radius = P.radius;
        h.n = n;
        h.n_tiles = n_tiles;
        *hdr = h;
    }
    if (idx < TABLE_ENTRIES) gauss_table_entry(P, idx, tab);
}

// Binning aggregated per block of BIN_THREADS emitters in LDS (one device
// atomic per block and tile; a window of BIN_WIN tiles from the first tile of
// the block's first frame).  A window of 2R + 1 <= 15 pixels spans at most
// 2 x 2 tiles.
// EPT emitters per thread (coalesced: emitter base + e * BIN_THREADS + tid):
// with one frame of many tiles (1 M emitters on 2048^2) a block's emitters
// spread over most of its window, and more of them per block means fewer
// device atomics per entry.
__device__ inline int gauss_window_base(const int* __restrict__ frame_id,
                                        const smx_fov_t* __restrict__ frames, int nframes, int n,
                                        int ept) {
    const long long i0 = (long long)blockIdx.x * smx::BIN_THREADS * ept;
    const int f0 = i0 < n ? frame_of(frame_id, (int)i0) : -1;
    return f0 >= 0 && f0 < nframes ? frames[f0].tile_off : 0;
}

// Amplitude of an emitter: I exp(-|z| / depth scale) (I when z is NULL).
__device__ inline float gauss_amp(const float* __restrict__ I, const float* __restrict__ z,
                                  float inv_depth, long long i) {
    return I[i] * (z ? expf(-fabsf(z[i]) * inv_depth) : 1.0f);
}

// The count pass also packs each emitter's x, y and amplitude into one
// 16-byte record (pk[i]), which the forward tile passes gather per list entry:
// one line per entry instead of one per array when the lists hold emitters
// from all over the arrays (1,000,000 emitters on 2048^2, bfloat16 matrix
// cores: 0.159 -> 0.146 ms per frame with the binning).
template <int EPT>
__global__ void __launch_bounds__(smx::BIN_THREADS) gauss_bin_count_kernel(
        int n, const float* __restrict__ x, const float* __restrict__ y,
        const float* __restrict__ I, const float* __restrict__ z, float inv_depth,
        float4* __restrict__ pk,
        const int* __restrict__ frame_id, const smx_fov_t* __restrict__ frames, int nframes, int R,
        int n_tiles, int* __restrict__ tile_count, int cs) {
    __shared__ int s_hist[smx::BIN_WIN];
    smx::bin_hist_clear(s_hist);
    const int wbase = gauss_window_base(frame_id, frames, nframes, n, EPT);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const long long i = ((long long)blockIdx.x * EPT + e) * smx::BIN_THREADS + threadIdx.x;
        if (i < n) pk[i] = make_float4(x[i], y[i], gauss_amp(I, z, inv_depth, i), 0.0f);
        const int f = i < n ? frame_of(frame_id, (int)i) : -1;
        TileRange t;
        if (f >= 0 && f < nframes && window_tiles(frames[f], x[i], y[i], R, t)) {
            const smx_fov_t& F = frames[f];
            for (int r = t.r0; r <= t.r1; ++r)
                for (int c = t.c0; c <= t.c1; ++c) {
                    const int tile = F.tile_off + r * F.tiles_c + c;
                    if (tile >= 0 && tile < n_tiles)
                        smx::bin_count(s_hist, wbase, tile, tile_count, cs);
                }
        }
    }
    __syncthreads();
    smx::bin_count_flush(s_hist, wbase, tile_count, cs);
}

template <int EPT>
__global__ void __launch_bounds__(smx::BIN_THREADS) gauss_bin_fill_kernel(
        int n, const float* __restrict__ x, const float* __restrict__ y,
        const int* __restrict__ frame_id, const smx_fov_t* __restrict__ frames, int nframes, int R,
        int n_tiles, const int* __restrict__ tile_begin, int* __restrict__ tile_left, int cs,
        int* __restrict__ list, long long capacity, int* __restrict__ status) {
    __shared__ int s_hist[smx::BIN_WIN];
    smx::bin_hist_clear(s_hist);
    const int wbase = gauss_window_base(frame_id, frames, nframes, n, EPT);
    __syncthreads();
    int ent[EPT][4];   // window entries: tile offset << 16 | LDS rank (-1: none)
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
#pragma unroll
        for (int q = 0; q < 4; ++q) ent[e][q] = -1;
        const long long i = ((long long)blockIdx.x * EPT + e) * smx::BIN_THREADS + threadIdx.x;
        const int f = i < n ? frame_of(frame_id, (int)i) : -1;
        TileRange t;
        if (f >= 0 && f < nframes && window_tiles(frames[f], x[i], y[i], R, t)) {
            const smx_fov_t& F = frames[f];
            int q = 0;
            for (int r = t.r0; r <= t.r1; ++r)
                for (int c = t.c0; c <= t.c1; ++c) {
                    const int tile = F.tile_off + r * F.tiles_c + c;
                    if (tile < 0 || tile >= n_tiles) continue;
                    const unsigned d = (unsigned)(tile - wbase);
                    if (d < (unsigned)smx::BIN_WIN) {
                        // ranks stay below 2^16: at most EPT * BIN_THREADS entries per tile
                        ent[e][q & 3] = (int)(d << 16) | atomicAdd(&s_hist[d], 1);
                        ++q;
                    } else {
                        const int pos = atomicSub(&tile_left[(size_t)tile * cs], 1) - 1;
                        const long long at = (long long)tile_begin[tile] + pos;
                        if (at < capacity) list[at] = (int)i;
                        else atomicOr(status, 2);
                    }
                }
        }
    }
    __syncthreads();
    smx::bin_claim(s_hist, wbase, tile_left, cs);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const int i = (blockIdx.x * EPT + e) * smx::BIN_THREADS + threadIdx.x;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (ent[e][q] < 0) continue;
            const int d = ent[e][q] >> 16, rank = ent[e][q] & 0xffff;
            const long long at = (long long)tile_begin[wbase + d] + s_hist[d] + rank;
            if (at < capacity) list[at] = i;
            else atomicOr(status, 2);
        }
    }
}

// Whole frames per XCD (smx::xcd_tile) when there are at least as many
// frames as XCDs: a frame's tiles share emitters and dL/dmu halos in one L2
// (64 frames of 256^2: binned backward 0.068 -> 0.061 ms).
static int xcd_map(int nframes) { return nframes >= smx::NXCD ? 1 : 0; }

struct FwdArgs {
    const float* x;
    const float* y;
    const float* z;
    const float* I;
    const float4* pk;    // x, y, amplitude per emitter (the count pass packs them)
    const smx_fov_t* frames;
    int nframes;
    const int* tile_fov;
    const int* tile_begin;
    int* list;           // each tile's sorted list is written back in place
    long long capacity;
    smx_gauss_params_t P;
    float* mu;
    int* status;
    int xcd;             // 1: workgroups map to tiles by smx::xcd_tile
};

__global__ void __launch_bounds__(256) gauss_forward_kernel(FwdArgs A) {
    __shared__ int s_sort[PSORT_CAP];
    __shared__ float s_acc[4][QS * QS];
    __shared__ float s_g[FCHUNK][2][16];       // pixel factors along u (rows) and v
    __shared__ int2 s_lo[FCHUNK];              // window origin (frame pixels)
    __shared__ float2 s_xy[FCHUNK];
    __shared__ float s_amp[FCHUNK];
    __shared__ int4 s_qi[4][FCHUNK];           // per wave: rectangle in the quadrant
    __shared__ int s_part[4][FCHUNK * 4];      // per wave: emitter << 2 | part

    const int tile = A.xcd ? smx::xcd_tile(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    const int fr = A.tile_fov[tile];
    if (fr < 0 || fr >= A.nframes) return;
    const smx_fov_t F = A.frames[fr];
    const int local = tile - F.tile_off;
    const int tr = local / F.tiles_c, tc = local % F.tiles_c;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int qr0 = tr * TS + (wave >> 1) * QS, qc0 = tc * TS + (wave & 1) * QS;
    const int qh = min(QS, F.nw - qr0), qw = min(QS, F.nh - qc0);
    const bool active = qh > 0 && qw > 0;
    const int R = A.P.radius, W = 2 * R + 1;
    const float s = INV_SQRT2 / A.P.sigma;
    float* acc = s_acc[wave];
    for (int k = lane; k < QS * QS; k += 64) acc[k] = 0.0f;

    const int begin = A.tile_begin[tile];
    int count = A.tile_begin[tile + 1] - begin;
    if ((long long)begin + count > A.capacity) count = (int)max(0LL, A.capacity - begin);
    const int* list = A.list + begin;
    const bool sorted = count <= PSORT_CAP;
    if (count >= 1 && count <= 256) {
        // up to one entry per thread: rank sort by the workgroup
        smx::rank_sort(list, count, s_sort + 256, s_sort);
    } else if (count >= 1 && count <= WAVE_SORT_MAX) {
        if (wave == 0) wave_sort_list(list, count, lane, s_sort);
        __syncthreads();
    } else if (count >= 1 && sorted) {
        int np2 = 1;
        while (np2 < count) np2 <<= 1;
        for (int k = threadIdx.x; k < np2; k += blockDim.x) s_sort[k] = k < count ? list[k] : 0x7fffffff;
        __syncthreads();
        lds_bitonic_sort(s_sort, np2);
    }
    // the sorted list goes back in place for the binned backward
    if (sorted && count > 1)
        for (int k = threadIdx.x; k < count; k += blockDim.x) A.list[begin + k] = s_sort[k];

    for (int base = 0; base < count; base += FCHUNK) {
        const int nc = min(FCHUNK, count - base);
        if (threadIdx.x < nc) {
            const int e = sorted ? s_sort[base + threadIdx.x] : list[base + threadIdx.x];
            const float4 p = A.pk[e];   // x, y, amplitude (packed by the count pass)
            s_xy[threadIdx.x] = make_float2(p.x, p.y);
            s_lo[threadIdx.x] = make_int2((int)floorf(p.x) - R, (int)floorf(p.y) - R);
            s_amp[threadIdx.x] = p.z;
        }
        __syncthreads();
        for (int t = threadIdx.x; t < nc * 2 * W; t += blockDim.x) {
            const int c = t / (2 * W), r = t - c * (2 * W);
            const int ax = r >= W, k = ax ? r - W : r;
            const float pos = ax ? s_xy[c].y : s_xy[c].x;
            const int lo = ax ? s_lo[c].y : s_lo[c].x;
            // two erfc per entry and no three-way divergence: the erfc values
            // serve both one-sided cases, the straddling pixel (at most one
            // per axis and emitter) takes erf itself
            const float t0 = ((float)(lo + k) - pos) * s, t1 = ((float)(lo + k) + 1.0f - pos) * s;
            s_g[c][ax][k] = g_pixel_edges(t0, t1, erfcf(fabsf(t0)), erfcf(fabsf(t1)));
        }
        __syncthreads();
        if (active) {
            // rectangle of chunk emitter `lane` inside this quadrant:
            // {first pixel, W | magic << 5, area, (row offset | col offset << 8) in the window}
            int np = 0;
            if (lane < nc) {
                const int2 lo = s_lo[lane];
                const int r0 = max(lo.x, qr0), r1 = min(lo.x + W, qr0 + qh);
                const int c0 = max(lo.y, qc0), c1 = min(lo.y + W, qc0 + qw);
                if (r1 > r0 && c1 > c0) {
                    const int Wr = c1 - c0, area = (r1 - r0) * Wr;
                    np = (area + 63) >> 6;
                    s_qi[wave][lane] = make_int4((r0 - qr0) * QS + (c0 - qc0),
                                                 Wr | (((65536 + Wr - 1) / Wr) << 5), area,
                                                 (r0 - lo.x) | ((c0 - lo.y) << 8));
                }
            }
            int incl = np;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                int t = __shfl_up(incl, d);
                if (lane >= d) incl += t;
            }
            const int total = __shfl(incl, 63);
            for (int p = 0; p < np; ++p) s_part[wave][incl - np + p] = (lane << 2) | p;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            for (int b = 0; b < total; b += NSLOT) {
                float val[NSLOT];
                int pix[NSLOT];
                // branch-free part decoding (24-bit multiplies; factor indices
                // clamped into their 16-entry rows, masked lanes add nothing),
                // so the LDS reads of the NSLOT parts issue back to back
#pragma unroll
                for (int q = 0; q < NSLOT; ++q) {
                    const int ent = s_part[wave][min(b + q, total - 1)], c = ent >> 2;
                    const int4 Q = s_qi[wave][c];
                    const int idx = (ent & 3) * 64 + lane;
                    const bool ok = (b + q < total) & (idx < Q.z);
                    const int Wr = Q.y & 31, magic = Q.y >> 5;
                    const int row = (int)(__umul24((unsigned)idx, (unsigned)magic) >> 16);
                    const int col = idx - __mul24(row, Wr);
                    const int ru = min((Q.w & 255) + row, 15), cv = min((Q.w >> 8) + col, 15);
                    val[q] = (s_amp[c] * s_g[c][0][ru]) * s_g[c][1][cv];
                    pix[q] = ok ? Q.x + row * QS + col : -1;
                }
#pragma unroll
                for (int q = 0; q < NSLOT; ++q)
                    if (pix[q] >= 0) lds_add(acc + pix[q], val[q]);
            }
        }
        __syncthreads();
    }
    if (active) {
        for (int k = lane; k < QS * QS; k += 64) {
            const int r = k / QS, c = k % QS;
            if (r >= qh || c >= qw) continue;
            const size_t o = (size_t)F.pix_off + (size_t)(qr0 + r) * F.nh + (qc0 + c);
            A.mu[o] = A.P.background + acc[k];
        }
    }
}

// dL/dmu of the Poisson NLL, 1 - k / mu, with the hardware reciprocal
// (v_rcp_f32, 1 ulp); every NLL path writes this value.
__device__ inline float nll_grad(float mu, float k) { return 1.0f - k * __builtin_amdgcn_rcpf(mu); }
// The fused tile pass's NLL term mu - k log(mu): the logarithm on the
// hardware log2 (v_log_f32, ~1e-7 relative), the term and its sums in
// float64.  gauss_nll_kernel evaluates the same term with a float64 log.
__device__ inline double nll_term(float mu, float k) {
    const float lg = 0.69314718055994531f * __builtin_amdgcn_logf(mu);
    return (double)mu - (double)k * (double)lg;
}

// Poisson NLL per frame, sum over pixels of mu - k log(mu) (float64), and
// dL/dmu (nll_grad).  Block b of frame f reduces pixels [b * per, (b+1) * per).
__global__ void __launch_bounds__(256) gauss_nll_kernel(const float* __restrict__ mu,
                                                        const float* __restrict__ k,
                                                        const smx_fov_t* __restrict__ frames,
                                                        int per, double* __restrict__ partial,
                                                        float* __restrict__ grad) {
    __shared__ double s_red[256];
    const int f = blockIdx.y, nblk = gridDim.x;
    const smx_fov_t F = frames[f];
    const long long npix = (long long)F.nw * F.nh;
    const long long start = (long long)blockIdx.x * per;
    const long long stop = min(start + (long long)per, npix);
    double sum = 0.0;
    for (long long p = start + threadIdx.x; p < stop; p += blockDim.x) {
        const size_t o = (size_t)F.pix_off + (size_t)p;
        const float m = mu[o], kk = k[o];
        const double md = (double)(m > 0.0f ? m : 1.0e-30f);
        sum += md - (double)kk * log(md);
        if (grad) grad[o] = nll_grad(m > 0.0f ? m : 1.0e-30f, kk);
    }
    s_red[threadIdx.x] = sum;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s_red[threadIdx.x] += s_red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[(size_t)f * nblk + blockIdx.x] = s_red[0];
}

__global__ void gauss_nll_finish_kernel(const double* __restrict__ partial, int nblk, int nframes,
                                        double* __restrict__ loss) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nframes) return;
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += partial[(size_t)f * nblk + b];
    loss[f] = s;
}

struct BwdArgs {
    int n;
    const float* x;
    const float* y;
    const float* z;
    const float* I;
    const int* frame_id;
    const smx_fov_t* frames;
    int nframes;
    smx_gauss_params_t P;
    const float* grad_mu;
    float* gx;
    float* gy;
    float* gz;
    float* gI;
    double* gsig_part;   // per emitter: dL/dsigma contribution (may be NULL)
    const float2* table; // window factors by phase (gauss_table_kernel)
};

// Writes gI, gx, gy, gz of emitter e (e >= 0) from the reduced sums pI, px,
// py and returns Ia * ps (ps: dmu/dsigma . g).
__device__ inline float emitter_outputs(float pI, float px, float py, float ps, int e, bool valid,
                                        float ze, float Ie, float inv_depth, float* gx, float* gy,
                                        float* gz, float* gI) {
    const float a = valid ? expf(-fabsf(ze) * inv_depth) : 0.0f;
    const float Ia = Ie * a;
    if (e >= 0) {
        const float sgn = ze > 0.0f ? 1.0f : (ze < 0.0f ? -1.0f : 0.0f);
        if (gI) gI[e] = a * pI;
        if (gx) gx[e] = Ia * px;
        if (gy) gy[e] = Ia * py;
        if (gz) gz[e] = -sgn * inv_depth * Ia * pI;
    }
    return Ia * ps;
}

// Gradients of one emitter from the five separable sums (a 16-lane group,
// lane j holding window column j; lane 0 writes): column j's terms t1 gy,
// t2 gy, t1 dgy and t3 gy + t1 sgy are summed by the xor-shuffle tree over
// strides 8, 4, 2, 1.  Returns Ia * dmu/dsigma . g (lane 0).
__device__ inline float group_outputs(float t1, float t2, float t3, float gyc, float dgyc,
                                      float sgyc, int j, int e, bool valid, float ze, float Ie,
                                      float inv_depth, float* gx, float* gy, float* gz,
                                      float* gI) {
    float pI = t1 * gyc, px = t2 * gyc, py = t1 * dgyc, ps = t3 * gyc + t1 * sgyc;
#pragma unroll
    for (int m = 8; m > 0; m >>= 1) {
        pI += __shfl_xor(pI, m);
        px += __shfl_xor(px, m);
        py += __shfl_xor(py, m);
        ps += __shfl_xor(ps, m);
    }
    return emitter_outputs(pI, px, py, ps, j == 0 ? e : -1, valid, ze, Ie, inv_depth, gx, gy, gz,
                           gI);
}

__global__ void __launch_bounds__(256) gauss_backward_kernel(BwdArgs A) {
    __shared__ float4 s_rowf[16][WMAX];   // per group, window row k: (G, dG/dx, dG/dsigma)
    const int e = blockIdx.x * 16 + (threadIdx.x >> 4);
    const int j = threadIdx.x & 15;
    const int R = A.P.radius, W = 2 * R + 1;
    bool valid = e < A.n;
    const int f = valid ? frame_of(A.frame_id, e) : 0;
    valid = valid && f >= 0 && f < A.nframes;
    smx_fov_t F;
    F.nw = F.nh = 0;
    F.pix_off = 0;
    float xe = 0.0f, ye = 0.0f, ze = 0.0f, Ie = 0.0f;
    if (valid) {
        F = A.frames[f];
        xe = A.x[e];
        ye = A.y[e];
        ze = A.z[e];
        Ie = A.I[e];
        valid = finite_coord(xe) && finite_coord(ye);
    }
    const int lu = valid ? (int)floorf(xe) - R : 0, lv = valid ? (int)floorf(ye) - R : 0;
    // this lane's window row factors (row lu + j) and column factors (col lv + j):
    // G, dG/dx and dG/dsigma from the phase table
    float4 fx = make_float4(0.0f, 0.0f, 0.0f, 0.0f), fy = fx;
    if (valid && j < W) {
        fx = table_factor(A.table, xe - floorf(xe), j);
        fy = table_factor(A.table, ye - floorf(ye), j);
    }
    // the group's row factors go through LDS: one 16-byte broadcast read per
    // row instead of three cross-lane shuffles
    float4* rowf = s_rowf[threadIdx.x >> 4];
    if (j < W) rowf[j] = fx;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int col = lv + j;
    const bool col_in = valid && j < W && col >= 0 && col < F.nh;
    float t1 = 0.0f, t2 = 0.0f, t3 = 0.0f;
    for (int k = 0; k < W; ++k) {
        const float4 fk = rowf[k];
        const int row = lu + k;
        float g = 0.0f;
        if (col_in && row >= 0 && row < F.nw)
            g = A.grad_mu[(size_t)F.pix_off + (size_t)row * F.nh + col];
        t1 = fmaf(g, fk.x, t1);
        t2 = fmaf(g, fk.y, t2);
        t3 = fmaf(g, fk.z, t3);
    }
    const float ps = group_outputs(t1, t2, t3, fy.x, fy.y, fy.z, j, e < A.n ? e : -1, valid, ze,
                                   Ie, A.P.inv_depth, A.gx, A.gy, A.gz, A.gI);
    if (j == 0 && e < A.n && A.gsig_part) A.gsig_part[e] = (double)ps;
}

// Binned backward: one workgroup per tile of the forward's binning.  The
// tile's window of dL/dmu (the tile plus R pixels on every side, zero off
// the frame) is staged in LDS once; the tile's list (as the forward left it:
// in emitter order when it was sorted) is filtered to the emitters this tile
// owns (owner_tile), and each owned emitter is reduced by a 16-lane group
// exactly as gauss_backward_kernel does, reading LDS instead of the image.
// Per tile the dL/dsigma terms are summed in list order with a fixed tree.
// Workgroups past n_tiles write zero gradients for the emitters no tile owns.
constexpr int HALO_MAX = TS + 2 * 7;
constexpr int BT_THREADS = 256;
constexpr int BT_GROUPS = BT_THREADS / 4;   // emitters per round (four lanes each)

struct BwdTileArgs {
    int n;
    const float* x;
    const float* y;
    const float* z;
    const float* I;
    const int* frame_id;
    const smx_fov_t* frames;
    int nframes;
    int n_tiles;
    const int* tile_fov;
    const int* tile_begin;
    const int* list;
    long long capacity;
    smx_gauss_params_t P;
    const float2* table;
    const GaussHeader* hdr;
    const float* grad_mu;
    float* gx;
    float* gy;
    float* gz;
    float* gI;
    double* tile_gsig;
    int* status;
    int xcd;             // 1: workgroups map to tiles by smx::xcd_tile
};

__global__ void __launch_bounds__(BT_THREADS) gauss_backward_tile_kernel(BwdTileArgs A) {
    // the tile's window of dL/dmu with R zero guard rows above and below
    // (rows of an off-frame window centre past the window are off the frame)
    __shared__ float s_g[(HALO_MAX + 14) * HALO_MAX];
    __shared__ float4 s_em[BT_THREADS];       // owned emitters: x, y, z, I
    __shared__ int s_own[BT_THREADS];
    __shared__ double s_gs[BT_THREADS];
    // per emitter of a round, window row k: G, dG/dx, dG/dsigma
    // (a float4 per row: one 16-byte read; one padding entry per group puts
    // the 16 groups of a wave on distinct banks)
    __shared__ float4 s_rowf[BT_GROUPS][GB + 1];
    __shared__ int s_wcnt[BT_THREADS / 64];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int R = A.P.radius, W = 2 * R + 1;
    if ((int)blockIdx.x >= A.n_tiles) {
        const int sweep = gridDim.x - A.n_tiles;
        for (long long e = (long long)(blockIdx.x - A.n_tiles) * BT_THREADS + tid; e < A.n;
             e += (long long)sweep * BT_THREADS) {
            const int f = frame_of(A.frame_id, (int)e);
            const bool owned = f >= 0 && f < A.nframes &&
                               owner_tile(A.frames[f], A.x[e], A.y[e], R, A.n_tiles) >= 0;
            if (!owned) {
                if (A.gI) A.gI[e] = 0.0f;
                if (A.gx) A.gx[e] = 0.0f;
                if (A.gy) A.gy[e] = 0.0f;
                if (A.gz) A.gz[e] = 0.0f;
            }
        }
        return;
    }
    const int tile = A.xcd ? smx::xcd_tile(blockIdx.x, A.n_tiles) : (int)blockIdx.x;
    const int fr = A.tile_fov[tile];
    const int begin = A.tile_begin[tile];
    int count = A.tile_begin[tile + 1] - begin;
    if (tile == 0 && tid == 0) {
        const GaussHeader h = *A.hdr;
        if (h.sigma != A.P.sigma || h.radius != R || h.n != A.n || h.n_tiles != A.n_tiles)
            atomicOr(A.status, 16);
    }
    if (fr < 0 || fr >= A.nframes) {
        if (tid == 0) A.tile_gsig[tile] = 0.0;
        return;
    }
    const smx_fov_t F = A.frames[fr];
    const int local = tile - F.tile_off;
    const int tr = local / F.tiles_c, tc = local % F.tiles_c;
    const int HS = TS + 2 * R;
    const int hr0 = tr * TS - R, hc0 = tc * TS - R;
    // staged rows r < HS + 2 R: frame row hr0 - R + r (guard rows zero),
    // one row per wave and step, one column per lane; every load of a
    // thread is issued before the first store, so the staging waits for one
    // memory latency instead of one per row
    constexpr int SROWS = (HALO_MAX + 14 + BT_THREADS / 64 - 1) / (BT_THREADS / 64);
    float gv[SROWS];
#pragma unroll
    for (int i = 0; i < SROWS; ++i) {
        const int r = wave + i * (BT_THREADS / 64);
        const int pr = hr0 - R + r, pc = hc0 + lane;
        const bool ok = lane < HS && r >= R && r < HS + R && pr >= 0 && pr < F.nw && pc >= 0 &&
                        pc < F.nh;
        gv[i] = ok ? A.grad_mu[(size_t)F.pix_off + (size_t)pr * F.nh + pc] : 0.0f;
    }
#pragma unroll
    for (int i = 0; i < SROWS; ++i) {
        const int r = wave + i * (BT_THREADS / 64);
        if (lane < HS && r < HS + 2 * R) s_g[r * HS + lane] = gv[i];
    }
    if ((long long)begin + count > A.capacity) count = (int)max(0LL, A.capacity - begin);
    const int* list = A.list + begin;
    // four lanes per emitter: lane q of group grp holds window slots
    // 4 q .. 4 q + 3 of both axes: the factors of its rows (into LDS) and of
    // its columns (registers) from two 16-byte table loads per axis and
    // quantity (one cache line per emitter, axis and quantity), then the
    // sums over all rows for its four columns
    const int grp = tid >> 2, q = tid & 3;
    const float4* tab4 = reinterpret_cast<const float4*>(A.table);
    double gsum = 0.0;
    for (int sb = 0; sb < count; sb += BT_THREADS) {
        const int ns = min(BT_THREADS, count - sb);
        int e = -1;
        bool own = false;
        float4 em = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (tid < ns) {
            e = list[sb + tid];
            em = make_float4(A.x[e], A.y[e], A.z[e], A.I[e]);
            own = owner_tile(F, em.x, em.y, R, A.n_tiles) == tile;
        }
        // order-preserving compaction of the owned entries
        const unsigned long long mask = __ballot(own);
        const int rank = __popcll(mask & ((1ull << lane) - 1ull));
        if (lane == 0) s_wcnt[wave] = __popcll(mask);
        __syncthreads();
        int off = 0;
        for (int w = 0; w < wave; ++w) off += s_wcnt[w];
        int nown = 0;
#pragma unroll
        for (int w = 0; w < BT_THREADS / 64; ++w) nown += s_wcnt[w];
        if (own) {
            s_own[off + rank] = e;
            s_em[off + rank] = em;
        }
        __syncthreads();
        for (int ob = 0; ob < nown; ob += BT_GROUPS) {
            if (ob + wave * (BT_GROUPS / 4) >= nown) continue;   // this wave's 16 emitters: none
            const int o = ob + grp;
            const bool valid = o < nown;
            const float4 v = s_em[min(o, nown - 1)];
            const int ee = valid ? s_own[o] : -1;
            const float fxl = floorf(v.x), fyl = floorf(v.y);
            const int lu = (int)fxl - R, lv = (int)fyl - R;
            float fu, fv;
            const int su = table_slot(v.x - fxl, fu) + 4 * q, sv = table_slot(v.y - fyl, fv) + 4 * q;
            float gy[4], dgy[4], sgy[4];
            float ru[3][4];
            int hcol[4];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float4* tu = tab4 + ((c * TABLE_ENTRIES + su) >> 1);
                const float4* tv = tab4 + ((c * TABLE_ENTRIES + sv) >> 1);
                const float4 u0 = tu[0], u1 = tu[1], v0 = tv[0], v1 = tv[1];
                ru[c][0] = fmaf(fu, u0.y, u0.x);
                ru[c][1] = fmaf(fu, u0.w, u0.z);
                ru[c][2] = fmaf(fu, u1.y, u1.x);
                ru[c][3] = fmaf(fu, u1.w, u1.z);
                const float rv[4] = {fmaf(fv, v0.y, v0.x), fmaf(fv, v0.w, v0.z),
                                     fmaf(fv, v1.y, v1.x), fmaf(fv, v1.w, v1.z)};
#pragma unroll
                for (int i = 0; i < 4; ++i) (c == 0 ? gy : c == 1 ? dgy : sgy)[i] = rv[i];
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (4 * q + i < W)
                    s_rowf[grp][4 * q + i] = make_float4(ru[0][i], ru[1][i], ru[2][i], 0.0f);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                // a column outside the staged block (off the frame) or past
                // the window reads a clamped cell and has zero factors
                const int jj = 4 * q + i, hc = lv + jj - hc0;
                const bool col_in = valid && jj < W && hc >= 0 && hc < HS;
                gy[i] = col_in ? gy[i] : 0.0f;
                dgy[i] = col_in ? dgy[i] : 0.0f;
                sgy[i] = col_in ? sgy[i] : 0.0f;
                hcol[i] = min(max(hc, 0), HS - 1);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // staged row of window row 0: lu - (hr0 - R) in [0, HS + R] for
            // every owned emitter
            const float* grow = s_g + (valid ? (lu - hr0 + R) : 0) * HS;
            float t1[4] = {0.0f, 0.0f, 0.0f, 0.0f}, t2[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            float t3[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int k = 0; k < WMAX; ++k) {
                if (k < W) {
                    const float4 fk = s_rowf[grp][k];
                    const float f0 = fk.x, f1 = fk.y, f2 = fk.z;
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const float g = grow[k * HS + hcol[i]];
                        t1[i] = fmaf(g, f0, t1[i]);
                        t2[i] = fmaf(g, f1, t2[i]);
                        t3[i] = fmaf(g, f2, t3[i]);
                    }
                }
            }
            float vI[4], vx[4], vy[4], vs[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                vI[i] = t1[i] * gy[i];
                vx[i] = t2[i] * gy[i];
                vy[i] = t1[i] * dgy[i];
                vs[i] = t3[i] * gy[i] + t1[i] * sgy[i];
            }
            float pI = (vI[0] + vI[1]) + (vI[2] + vI[3]);
            float px = (vx[0] + vx[1]) + (vx[2] + vx[3]);
            float py = (vy[0] + vy[1]) + (vy[2] + vy[3]);
            float ps = (vs[0] + vs[1]) + (vs[2] + vs[3]);
            // sums over the quad (lanes xor 1, then xor 2) on DPP lane moves
            pI += smx::dpp_f32<0xB1>(pI);
            px += smx::dpp_f32<0xB1>(px);
            py += smx::dpp_f32<0xB1>(py);
            ps += smx::dpp_f32<0xB1>(ps);
            pI += smx::dpp_f32<0x4E>(pI);
            px += smx::dpp_f32<0x4E>(px);
            py += smx::dpp_f32<0x4E>(py);
            ps += smx::dpp_f32<0x4E>(ps);
            const float gs = emitter_outputs(pI, px, py, ps, q == 0 ? ee : -1, valid, v.z, v.w,
                                             A.P.inv_depth, A.gx, A.gy, A.gz, A.gI);
            if (valid && q == 0) s_gs[o] = (double)gs;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        __syncthreads();
        if (wave == 0) {
            double v = 0.0;
#pragma unroll
            for (int r = 0; r < BT_THREADS / 64; ++r)
                if (r * 64 + lane < nown) v += s_gs[r * 64 + lane];
#pragma unroll
            for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
            gsum += v;
        }
        __syncthreads();
    }
    if (tid == 0) A.tile_gsig[tile] = gsum;
}

// Deterministic float64 sum (fixed assignment of values to threads and a
// fixed tree): block b sums in[b * per .. (b + 1) * per) into out[b].
constexpr int SUM_BLOCKS = 256;
__global__ void __launch_bounds__(256) sum_f64_kernel(const double* __restrict__ in, int n,
                                                      int per, double* __restrict__ out) {
    __shared__ double s_red[256];
    const int start = blockIdx.x * per, stop = min(n, start + per);
    double s = 0.0;
    for (int i = start + threadIdx.x; i < stop; i += blockDim.x) s += in[i];
    s_red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s_red[threadIdx.x] += s_red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = s_red[0];
}

// --------------------------------------------------------------------------
// Separable splat on the matrix cores.  Per 32 x 32 tile the image is the
// contraction  mu_tile = A B,  A[i][e] = amp_e G(i; x_e),  B[e][j] = G(j; y_e)
// over the tile's emitters e (zero outside each emitter's window): the tile's
// sorted list is staged SG_STAGE emitters at a time (window origin, phases,
// amplitude in LDS), each of the SG_WAVES waves takes every SG_WAVES-th chunk
// of 16 emitters, builds its factors from the phase table and adds them into
// its own 32 x 32 MFMA accumulator, and the waves' accumulators are summed in
// a fixed order at the end.  Inputs are rounded to the operand type (float32,
// float16 or bfloat16); accumulation is float32.
// --------------------------------------------------------------------------
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int SG_WAVES = 4;
constexpr int SG_THREADS = 64 * SG_WAVES;
constexpr int SG_STAGE = SG_THREADS;
constexpr int SG_SORT_CAP = PSORT_CAP;   // tile lists sorted in LDS (longer: after binning)

// Factor rows of a chunk emitter: [0] = 0, [1 .. W] = amp * G along the rows,
// [W + 1 ..] = 0, then the same 18 slots along the columns; an MFMA operand
// is row[clamp(pixel - window origin, -1, W) + 1], zero outside the window
// with no mask.
constexpr int SG_AXIS = 18;
constexpr int SG_ROW = 2 * SG_AXIS;

struct SepArgs {
    const float* x;
    const float* y;
    const float* z;      // may be NULL (amplitude I)
    const float* I;
    const float4* pk;    // x, y, amplitude per emitter (the count pass packs them)
    const float2* tab;   // G block of the phase table
    float inv_depth;
    int R;
    int W;
    const smx_fov_t* frames;
    int nframes;
    const int* tile_fov;
    const int* tile_begin;
    int* list;           // each tile's sorted list is written back in place
    long long capacity;
    float background;
    float* mu;           // may be NULL when the NLL is fused (NLL)
    int* status;
    const float* k;      // NLL: observed counts per pixel
    float* grad_mu;      // NLL: dL/dmu = 1 - k / mu (may be NULL)
    double* tile_nll;    // NLL: sum over the tile of mu - k log(mu)
    int xcd;             // 1: workgroups map to tiles by smx::xcd_tile
};

struct SepLds {
    int sort[SG_SORT_CAP];                // the tile's list in index order
    union alignas(16) {
        int raw[SG_WAVES * 256];          // sorted runs of the list (merge_rank_sort)
        float red[16 * 64];               // the waves' accumulators, summed in turn
    } u;
    int2 lo[SG_STAGE];                    // window origin (frame pixels)
    float2 ph[SG_STAGE];                  // sub-pixel phases
    float amp[SG_STAGE];
    float fac[SG_WAVES][16][SG_ROW];       // per wave: one chunk's factors (padded rows)
    double nll[SG_WAVES];
};

template <typename T> struct Frag8 { typedef T type __attribute__((ext_vector_type(8))); };

// DT: operand type, 0: float32 (32x32x2), 1: float16 (32x32x16), 2: bfloat16
// (32x32x16).  NLL: the Poisson NLL of the tile (nll_term, summed in float64)
// and its image gradient (nll_grad) in the epilogue, from the accumulators,
// instead of a pass over the written image.
template <int DT, bool NLL>
__global__ void __launch_bounds__(SG_THREADS) __attribute__((amdgpu_waves_per_eu(6, 8)))
sep_gemm_kernel(SepArgs A) {
    __shared__ SepLds L;
    const int tile = A.xcd ? smx::xcd_tile(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    // independent loads first: frame, list bounds
    const int fr = A.tile_fov[tile];
    const int begin = A.tile_begin[tile];
    int count = A.tile_begin[tile + 1] - begin;
    if (fr < 0 || fr >= A.nframes) return;
    const smx_fov_t F = A.frames[fr];
    const int local = tile - F.tile_off;
    const int tr = local / F.tiles_c, tc = local % F.tiles_c;
    const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63;
    const int m = l & 31, half = l >> 5;
    const int row = tr * TS + m, col = tc * TS + m;   // A row / B column of this lane
    const int R = A.R, W = A.W;
    // this wave's outputs: rows 8 wave + 4 half + i (i < 4), column m
    const int pj = tc * TS + m;
    float kk[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (NLL && pj < F.nh) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int pi = tr * TS + 8 * wave + 4 * half + i;
            if (pi < F.nw) kk[i] = A.k[(size_t)F.pix_off + (size_t)pi * F.nh + pj];
        }
    }

    if ((long long)begin + count > A.capacity) count = (int)max(0LL, A.capacity - begin);
    int* list = A.list + begin;
    const bool sorted = count <= SG_SORT_CAP;
    if (count >= 2 && sorted) {
        // up to 1024 entries: every wave sorts a run in registers and the
        // runs are ranked against each other; longer: bitonic sort in LDS
        if (count <= SG_WAVES * 64) {
            smx::merge_rank_sort<1>(list, count, L.u.raw, L.sort);
        } else if (count <= SG_WAVES * 128) {
            smx::merge_rank_sort<2>(list, count, L.u.raw, L.sort);
        } else if (count <= SG_WAVES * 256) {
            smx::merge_rank_sort<4>(list, count, L.u.raw, L.sort);
        } else {
            int np2 = 1;
            while (np2 < count) np2 <<= 1;
            for (int k = tid; k < np2; k += SG_THREADS) L.sort[k] = k < count ? list[k] : 0x7fffffff;
            __syncthreads();
            lds_bitonic_sort(L.sort, np2);
        }
        // the sorted list goes back in place for the binned backward
        for (int k = tid; k < count; k += SG_THREADS) list[k] = L.sort[k];
    } else if (count == 1) {
        if (tid == 0) L.sort[0] = list[0];
        __syncthreads();
    }

    f32x16 acc;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.0f;
    float (*fac)[SG_ROW] = L.fac[wave];
    // leading zero of both axes of every row (never overwritten)
    if (l < 32) fac[l >> 1][(l & 1) * SG_AXIS] = 0.0f;
    // factor build: lane (k, ax, hb) = (l / 4, (l / 2) % 2, l % 2) writes
    // window pixels 8 hb .. 8 hb + 7 of axis ax of chunk emitter k: four
    // 16-byte table loads (value, slope pairs; zero pairs past the window)
    const int fk = l >> 2, fax = (l >> 1) & 1, fhb = l & 1;
    const float4* tab4 = reinterpret_cast<const float4*>(A.tab);
    for (int sb = 0; sb < count; sb += SG_STAGE) {
        const int ns = min(SG_STAGE, count - sb);
        if (tid < ns) {
            const int e = sorted ? L.sort[sb + tid] : list[sb + tid];
            const float4 p = A.pk[e];   // x, y, amplitude (packed by the count pass)
            const float xe = p.x, ye = p.y;
            const float fx = floorf(xe), fy = floorf(ye);
            const bool fin = finite_coord(xe) && finite_coord(ye);
            L.lo[tid] = fin ? make_int2((int)fx - R, (int)fy - R) : make_int2(0, 0);
            L.ph[tid] = fin ? make_float2(xe - fx, ye - fy) : make_float2(0.0f, 0.0f);
            L.amp[tid] = fin ? p.z : 0.0f;
        }
        __syncthreads();
        const int nch = (ns + 15) >> 4;
        // this lane's table pairs, interpolation weight and scale (amplitude,
        // 1, or 0 past the chunk) of chunk c, loaded one chunk ahead of use
        float4 gq[4];
        float gf = 0.0f, gs = 0.0f;
        auto load_g = [&](int c) {
            const int kb = c * 16, nk = min(16, ns - kb);
            const int k = min(fk, nk - 1);
            const float2 ph = L.ph[kb + k];
            float f;
            const int slot = table_slot(fax ? ph.y : ph.x, f) + fhb * 8;
            const float4* src = tab4 + (slot >> 1);
#pragma unroll
            for (int i = 0; i < 4; ++i) gq[i] = src[i];
            gf = f;
            gs = fk < nk ? (fax ? 1.0f : L.amp[kb + k]) : 0.0f;
        };
        if (wave < nch) load_g(wave);
        for (int c = wave; c < nch; c += SG_WAVES) {
            const int kb = c * 16;
            float v[8];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                v[2 * i] = fmaf(gf, gq[i].y, gq[i].x) * gs;
                v[2 * i + 1] = fmaf(gf, gq[i].w, gq[i].z) * gs;
            }
            if (c + SG_WAVES < nch) load_g(c + SG_WAVES);
            float* dst = &fac[fk][fax * SG_AXIS + 1 + fhb * 8];
#pragma unroll
            for (int i = 0; i < 8; ++i) dst[i] = v[i];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // operands: row[clamp(offset, -1, W) + 1] (rows past the chunk are zero)
            if (DT == 0) {
                int2 o[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) o[j] = L.lo[kb + 2 * j + half];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int k = 2 * j + half;
                    const int ia = min(max(row - o[j].x, -1), W) + 1;
                    const int ib = min(max(col - o[j].y, -1), W) + 1;
                    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(fac[k][ia], fac[k][SG_AXIS + ib],
                                                               acc, 0, 0, 0);
                }
            } else {
                typedef typename std::conditional<DT == 1, _Float16, __bf16>::type E;
                typedef typename Frag8<E>::type V8;
                V8 av, bv;
                int2 o[8];
#pragma unroll
                for (int t = 0; t < 8; ++t) o[t] = L.lo[kb + half * 8 + t];
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    const int k = half * 8 + t;
                    const int ia = min(max(row - o[t].x, -1), W) + 1;
                    const int ib = min(max(col - o[t].y, -1), W) + 1;
                    av[t] = (E)fac[k][ia];
                    bv[t] = (E)fac[k][SG_AXIS + ib];
                }
                if constexpr (DT == 1) acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(av, bv, acc, 0, 0, 0);
                else acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc, 0, 0, 0);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        __syncthreads();
    }
    // acc[i] holds row 8 (i / 4) + 4 half + i % 4, column m of the tile.  The
    // waves' accumulators are added into red[] in wave order, then wave w
    // finishes rows 8 w .. 8 w + 7.
#pragma unroll
    for (int w = 0; w < SG_WAVES; ++w) {
        if (wave == w) {
#pragma unroll
            for (int i = 0; i < 16; ++i)
                L.u.red[i * 64 + l] = w == 0 ? acc[i] : L.u.red[i * 64 + l] + acc[i];
        }
        __syncthreads();
    }
    double part = 0.0;
    if (pj < F.nh) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int pi = tr * TS + 8 * wave + 4 * half + i;
            if (pi >= F.nw) continue;
            const float v = L.u.red[(4 * wave + i) * 64 + l];
            const size_t o = (size_t)F.pix_off + (size_t)pi * F.nh + pj;
            const float mv = A.background + v;
            if (A.mu) A.mu[o] = mv;
            if (NLL) {
                const float mc = mv > 0.0f ? mv : 1.0e-30f;
                part += nll_term(mc, kk[i]);
                if (A.grad_mu) A.grad_mu[o] = nll_grad(mc, kk[i]);
            }
        }
    }
    if (NLL) {
        // fixed reduction trees (over the wave, then the waves in order): the
        // tile sum is deterministic
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
        if (l == 0) L.nll[wave] = part;
        __syncthreads();
        if (tid == 0) {
            double t = L.nll[0];
#pragma unroll
            for (int w = 1; w < SG_WAVES; ++w) t += L.nll[w];
            A.tile_nll[tile] = t;
        }
    }
}

// loss[f] = sum of frame f's tile sums (one wave per frame: lanes take the
// tiles in a fixed stride, then a fixed shuffle tree).
__global__ void __launch_bounds__(256) gauss_tile_nll_finish_kernel(
        const smx_fov_t* __restrict__ frames, int nframes, const double* __restrict__ tile_nll,
        double* __restrict__ loss) {
    const int f = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (f >= nframes) return;
    const smx_fov_t F = frames[f];
    const int nt = F.tiles_r * F.tiles_c;
    double s = 0.0;
    for (int t = lane; t < nt; t += 64) s += tile_nll[F.tile_off + t];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) loss[f] = s;
}

struct GaussWorkspace {
    int* tile_count;     // (n_tiles + 1) x COUNT_STRIDE (padded counters)
    int* tile_dense;     // n_tiles + 1
    int* tile_begin;     // n_tiles + 1
    int* tile_fov;       // n_tiles
    int* list;           // capacity
    float2* table;       // TABLE_PAIRS window factors by phase
    GaussHeader* hdr;    // what the binning and table were built for
    double* tile_nll;    // n_tiles: fused NLL per tile
    double* tile_gsig;   // n_tiles: binned backward's dL/dsigma per tile
    int* long_tiles;     // n_tiles + 1: tiles with lists too long to sort in LDS
    float4* pk;          // n: x, y, amplitude per emitter (written by the count pass)
    void* cub_tmp;
    size_t cub_bytes;
};

size_t gauss_layout(int n, int n_tiles, long long capacity, char* base, GaussWorkspace* ws) {
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += align256(bytes); return o; };
    size_t o_count = take(sizeof(int) * (size_t)(n_tiles + 1) * smx::COUNT_STRIDE);
    size_t o_dense = take(sizeof(int) * (n_tiles + 1));
    size_t o_begin = take(sizeof(int) * (n_tiles + 1));
    size_t o_fov = take(sizeof(int) * n_tiles);
    size_t o_list = take(sizeof(int) * (size_t)(capacity > 0 ? capacity : 1));
    size_t o_tab = take(sizeof(float2) * (size_t)TABLE_PAIRS);
    size_t o_hdr = take(sizeof(GaussHeader));
    size_t o_tnll = take(sizeof(double) * (size_t)n_tiles);
    size_t o_tgs = take(sizeof(double) * (size_t)n_tiles);
    size_t o_long = take(sizeof(int) * (size_t)(n_tiles + 1));
    size_t o_pk = take(sizeof(float4) * (size_t)(n > 0 ? n : 1));
    size_t cb = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, cb, (int*)nullptr, (int*)nullptr, n_tiles + 1);
    size_t o_cub = take(cb);
    if (ws && base) {
        ws->tile_count = (int*)(base + o_count);
        ws->tile_dense = (int*)(base + o_dense);
        ws->tile_begin = (int*)(base + o_begin);
        ws->tile_fov = (int*)(base + o_fov);
        ws->list = (int*)(base + o_list);
        ws->table = (float2*)(base + o_tab);
        ws->hdr = (GaussHeader*)(base + o_hdr);
        ws->tile_nll = (double*)(base + o_tnll);
        ws->tile_gsig = (double*)(base + o_tgs);
        ws->long_tiles = (int*)(base + o_long);
        ws->pk = (float4*)(base + o_pk);
        ws->cub_tmp = base + o_cub;
        ws->cub_bytes = cb;
    }
    return off;
}

bool params_ok(const smx_gauss_params_t* P) {
    return P && P->sigma > 0.0f && P->radius >= 1 && 2 * P->radius + 1 <= WMAX &&
           P->inv_depth >= 0.0f;
}

constexpr int NLL_PER = 4096;

// backward workspace: per-emitter dL/dsigma terms and their partial sums
// (float64), then the phase table
size_t backward_table_offset(int n) {
    return align256(sizeof(double) * ((size_t)(n > 0 ? n : 1) + SUM_BLOCKS));
}

}  // namespace

extern "C" {

size_t smx_gauss_workspace_bytes(int n, int n_tiles, long long capacity) {
    if (n < 0 || n_tiles <= 0 || capacity < 0) return 0;
    return gauss_layout(n, n_tiles, capacity, nullptr, nullptr);
}

size_t smx_gauss_backward_workspace_bytes(int n) {
    if (n < 0) return 0;
    return backward_table_offset(n) + sizeof(float2) * (size_t)TABLE_PAIRS;
}

}  // extern "C"

namespace {

// Tile binning shared by every forward form: init (counters, tile -> frame
// map, phase table) -> count -> scan -> fill.
int gauss_bin(int n, const float* x, const float* y, const float* z, const float* I,
              const int* frame_id,
              const smx_fov_t* frames_dev, int nframes, int n_tiles,
              const smx_gauss_params_t& P, long long capacity, GaussWorkspace& ws,
              int* status_dev, hipStream_t stream) {
    const int R = P.radius;
    // four emitters per thread for large launches (denser LDS windows)
    const int ept = n >= (1 << 19) ? 4 : 1;
    const int nbb = (int)(((long long)n + (long long)smx::BIN_THREADS * ept - 1) /
                          ((long long)smx::BIN_THREADS * ept));
    // dense counters when the blocks' LDS windows hold every tile; beyond
    // that, padded ones only when a tile holds hundreds of entries: a 13 px
    // window enters ~2 tiles, and the single-workgroup scan reads padded
    // counters one cache line each
    const int cs = smx::agg_count_stride(2LL * n, n_tiles, 256);
    const int tile_blocks = (n_tiles + 1 + 255) / 256;
    const int table_blocks = (TABLE_ENTRIES + 255) / 256;
    hipLaunchKernelGGL(gauss_init_kernel, dim3(tile_blocks + table_blocks), dim3(256), 0, stream,
                       nframes, frames_dev, n_tiles, ws.tile_count, cs, ws.tile_fov, tile_blocks,
                       P, n, ws.table, ws.hdr, ws.long_tiles);
    SMX_CHECK_LAUNCH("gauss_init_kernel");
    if (n > 0) {
        if (ept == 4)
            hipLaunchKernelGGL(gauss_bin_count_kernel<4>, dim3(nbb), dim3(smx::BIN_THREADS), 0,
                               stream, n, x, y, I, z, P.inv_depth, ws.pk, frame_id, frames_dev,
                               nframes, R, n_tiles, ws.tile_count, cs);
        else
            hipLaunchKernelGGL(gauss_bin_count_kernel<1>, dim3(nbb), dim3(smx::BIN_THREADS), 0,
                               stream, n, x, y, I, z, P.inv_depth, ws.pk, frame_id, frames_dev,
                               nframes, R, n_tiles, ws.tile_count, cs);
        SMX_CHECK_LAUNCH("gauss_bin_count_kernel");
    }
    if (n_tiles + 1 <= smx::SCAN_MAX) {
        // one workgroup scans the counters in place of compact + device scan
        hipLaunchKernelGGL(smx::scan_counts_kernel, dim3(1), dim3(smx::SCAN_THREADS), 0, stream,
                           ws.tile_count, cs, n_tiles + 1, ws.tile_begin, PSORT_CAP,
                           ws.long_tiles);
        SMX_CHECK_LAUNCH("scan_counts_kernel");
    } else {
        size_t cb = ws.cub_bytes;
        hipLaunchKernelGGL(smx::compact_counts_kernel, dim3((n_tiles + 1 + 255) / 256), dim3(256),
                           0, stream, ws.tile_count, cs, n_tiles + 1, ws.tile_dense, PSORT_CAP,
                           ws.long_tiles);
        SMX_CHECK_LAUNCH("compact_counts_kernel");
        if (smx::check_hip(hipcub::DeviceScan::ExclusiveSum(ws.cub_tmp, cb, ws.tile_dense,
                                                            ws.tile_begin, n_tiles + 1, stream),
                           "hipcub ExclusiveSum"))
            return SMX_ERUNTIME;
    }
    if (n > 0) {
        if (ept == 4)
            hipLaunchKernelGGL(gauss_bin_fill_kernel<4>, dim3(nbb), dim3(smx::BIN_THREADS), 0,
                               stream, n, x, y, frame_id, frames_dev, nframes, R, n_tiles,
                               ws.tile_begin, ws.tile_count, cs, ws.list, capacity, status_dev);
        else
            hipLaunchKernelGGL(gauss_bin_fill_kernel<1>, dim3(nbb), dim3(smx::BIN_THREADS), 0,
                               stream, n, x, y, frame_id, frames_dev, nframes, R, n_tiles,
                               ws.tile_begin, ws.tile_count, cs, ws.list, capacity, status_dev);
        SMX_CHECK_LAUNCH("gauss_bin_fill_kernel");
        hipLaunchKernelGGL(smx::sort_long_lists_kernel, dim3(smx::LONG_SORT_BLOCKS),
                           dim3(smx::LONG_SORT_THREADS), 0, stream, ws.long_tiles, ws.tile_begin,
                           ws.list, capacity, status_dev);
        SMX_CHECK_LAUNCH("sort_long_lists_kernel");
    }
    return SMX_OK;
}

int forward_checks(const char* who, int n, const float* x, const float* y, const float* z,
                   const float* I, const smx_fov_t* frames_dev, int nframes, int n_tiles,
                   const smx_gauss_params_t* params, long long capacity, const float* mu,
                   void* workspace, size_t workspace_bytes, const int* status_dev,
                   bool z_required) {
    if (n < 0 || nframes <= 0 || n_tiles <= 0 || !frames_dev || !mu || !workspace ||
        !status_dev || !params_ok(params) || capacity < 0 ||
        (n > 0 && (!x || !y || !I || (z_required && !z)))) {
        smx::set_error("%s: invalid arguments (n=%d, nframes=%d, n_tiles=%d)", who, n, nframes,
                       n_tiles);
        return SMX_EVALUE;
    }
    size_t need = gauss_layout(n, n_tiles, capacity, nullptr, nullptr);
    if (workspace_bytes < need) {
        smx::set_error("%s: workspace too small (%zu < %zu bytes)", who, workspace_bytes, need);
        return SMX_EVALUE;
    }
    return SMX_OK;
}

}  // namespace

extern "C" {

int smx_gauss_forward(int n, const float* x, const float* y, const float* z, const float* I,
                      const int* frame_id, const smx_fov_t* frames_dev, int nframes, int n_tiles,
                      const smx_gauss_params_t* params, long long capacity, float* mu,
                      void* workspace, size_t workspace_bytes, int* status_dev,
                      hipStream_t stream) {
    int rc = forward_checks("smx_gauss_forward", n, x, y, z, I, frames_dev, nframes, n_tiles,
                            params, capacity, mu, workspace, workspace_bytes, status_dev, true);
    if (rc) return rc;
    GaussWorkspace ws;
    gauss_layout(n, n_tiles, capacity, (char*)workspace, &ws);
    rc = gauss_bin(n, x, y, z, I, frame_id, frames_dev, nframes, n_tiles, *params, capacity, ws,
                   status_dev, stream);
    if (rc) return rc;
    FwdArgs A;
    A.x = x; A.y = y; A.z = z; A.I = I; A.pk = ws.pk; A.frames = frames_dev; A.nframes = nframes;
    A.tile_fov = ws.tile_fov; A.tile_begin = ws.tile_begin; A.list = ws.list;
    A.capacity = capacity; A.P = *params; A.mu = mu; A.status = status_dev; A.xcd = xcd_map(nframes);
    hipLaunchKernelGGL(gauss_forward_kernel, dim3(n_tiles), dim3(256), 0, stream, A);
    SMX_CHECK_LAUNCH("gauss_forward_kernel");
    return SMX_OK;
}

}  // extern "C"

namespace {

// Binning, phase table, factors and the matrix-core tile pass (with the NLL
// epilogue when k is given).
int sep_gemm_run(int n, const float* x, const float* y, const float* z, const float* I,
                 const int* frame_id, const smx_fov_t* frames_dev, int nframes, int n_tiles,
                 const smx_gauss_params_t* params, int dtype, long long capacity, float* mu,
                 const float* k, float* grad_mu, double* loss, void* workspace,
                 size_t workspace_bytes, int* status_dev, hipStream_t stream) {
    if (dtype < 0 || dtype > 2) {
        smx::set_error("smx_sep_gemm: dtype must be 0 (float32), 1 (float16) or 2 (bfloat16)");
        return SMX_EVALUE;
    }
    GaussWorkspace ws;
    gauss_layout(n, n_tiles, capacity, (char*)workspace, &ws);
    int rc = gauss_bin(n, x, y, z, I, frame_id, frames_dev, nframes, n_tiles, *params, capacity,
                       ws, status_dev, stream);
    if (rc) return rc;
    SepArgs A;
    A.x = x; A.y = y; A.z = z; A.I = I; A.pk = ws.pk; A.tab = ws.table;
    A.inv_depth = params->inv_depth;
    A.R = params->radius; A.W = 2 * params->radius + 1;
    A.frames = frames_dev; A.nframes = nframes;
    A.tile_fov = ws.tile_fov; A.tile_begin = ws.tile_begin; A.list = ws.list;
    A.capacity = capacity; A.background = params->background;
    A.mu = mu; A.status = status_dev;
    A.k = k; A.grad_mu = grad_mu; A.tile_nll = ws.tile_nll; A.xcd = xcd_map(nframes);
    auto launch = [&](auto nll) {
        constexpr bool L = decltype(nll)::value;
        const dim3 grid(n_tiles), block(SG_THREADS);
        if (dtype == 0) hipLaunchKernelGGL((sep_gemm_kernel<0, L>), grid, block, 0, stream, A);
        else if (dtype == 1) hipLaunchKernelGGL((sep_gemm_kernel<1, L>), grid, block, 0, stream, A);
        else hipLaunchKernelGGL((sep_gemm_kernel<2, L>), grid, block, 0, stream, A);
    };
    if (k) launch(std::true_type());
    else launch(std::false_type());
    SMX_CHECK_LAUNCH("sep_gemm_kernel");
    if (k) {
        hipLaunchKernelGGL(gauss_tile_nll_finish_kernel, dim3((nframes + 3) / 4), dim3(256), 0,
                           stream, frames_dev, nframes, ws.tile_nll, loss);
        SMX_CHECK_LAUNCH("gauss_tile_nll_finish_kernel");
    }
    return SMX_OK;
}

}  // namespace

extern "C" {

int smx_sep_gemm(int n, const float* x, const float* y, const float* z, const float* I,
                 const int* frame_id, const smx_fov_t* frames_dev, int nframes, int n_tiles,
                 const smx_gauss_params_t* params, int dtype, long long capacity, float* mu,
                 void* workspace, size_t workspace_bytes, int* status_dev, hipStream_t stream) {
    int rc = forward_checks("smx_sep_gemm", n, x, y, z, I, frames_dev, nframes, n_tiles, params,
                            capacity, mu, workspace, workspace_bytes, status_dev, false);
    if (rc) return rc;
    return sep_gemm_run(n, x, y, z, I, frame_id, frames_dev, nframes, n_tiles, params, dtype,
                        capacity, mu, nullptr, nullptr, nullptr, workspace, workspace_bytes,
                        status_dev, stream);
}

int smx_sep_gemm_nll(int n, const float* x, const float* y, const float* z, const float* I,
                     const int* frame_id, const smx_fov_t* frames_dev, int nframes, int n_tiles,
                     const smx_gauss_params_t* params, int dtype, long long capacity, float* mu,
                     const float* k, float* grad_mu, double* loss, void* workspace,
                     size_t workspace_bytes, int* status_dev, hipStream_t stream) {
    // mu may be NULL here: the loss and its gradient need no image in memory
    int rc = forward_checks("smx_sep_gemm_nll", n, x, y, z, I, frames_dev, nframes, n_tiles,
                            params, capacity, mu ? mu : (const float*)k, workspace,
                            workspace_bytes, status_dev, false);
    if (rc) return rc;
    if (!k || !loss) {
        smx::set_error("smx_sep_gemm_nll: counts k and the loss buffer are required");
        return SMX_EVALUE;
    }
    return sep_gemm_run(n, x, y, z, I, frame_id, frames_dev, nframes, n_tiles, params, dtype,
                        capacity, mu, k, grad_mu, loss, workspace, workspace_bytes, status_dev,
                        stream);
}

size_t smx_gauss_nll_workspace_bytes(int nframes, int max_frame_pixels) {
    if (nframes <= 0 || max_frame_pixels <= 0) return 0;
    int nblk = (max_frame_pixels + NLL_PER - 1) / NLL_PER;
    return sizeof(double) * (size_t)nblk * nframes;
}

int smx_gauss_poisson_nll(const float* mu, const float* k, const smx_fov_t* frames_dev,
                          int nframes, int max_frame_pixels, double* loss, float* grad_mu,
                          void* workspace, size_t workspace_bytes, hipStream_t stream) {
    if (!mu || !k || !frames_dev || nframes <= 0 || max_frame_pixels <= 0 || !loss ||
        !workspace || workspace_bytes < smx_gauss_nll_workspace_bytes(nframes, max_frame_pixels)) {
        smx::set_error("smx_gauss_poisson_nll: invalid arguments (nframes=%d)", nframes);
        return SMX_EVALUE;
    }
    int nblk = (max_frame_pixels + NLL_PER - 1) / NLL_PER;
    hipLaunchKernelGGL(gauss_nll_kernel, dim3(nblk, nframes), dim3(256), 0, stream, mu, k,
                       frames_dev, NLL_PER, (double*)workspace, grad_mu);
    SMX_CHECK_LAUNCH("gauss_nll_kernel");
    hipLaunchKernelGGL(gauss_nll_finish_kernel, dim3((nframes + 255) / 256), dim3(256), 0, stream,
                       (const double*)workspace, nblk, nframes, loss);
    SMX_CHECK_LAUNCH("gauss_nll_finish_kernel");
    return SMX_OK;
}

int smx_gauss_backward(int n, const float* x, const float* y, const float* z, const float* I,
                       const int* frame_id, const smx_fov_t* frames_dev, int nframes,
                       const smx_gauss_params_t* params, const float* grad_mu, float* gx,
                       float* gy, float* gz, float* gI, double* gsigma, void* workspace,
                       size_t workspace_bytes, hipStream_t stream) {
    if (n < 0 || nframes <= 0 || !frames_dev || !grad_mu || !params_ok(params) ||
        (n > 0 && (!x || !y || !z || !I)) || !workspace ||
        workspace_bytes < smx_gauss_backward_workspace_bytes(n)) {
        smx::set_error("smx_gauss_backward: invalid arguments (n=%d, nframes=%d, workspace %zu "
                       "bytes)", n, nframes, workspace_bytes);
        return SMX_EVALUE;
    }
    float2* table = (float2*)((char*)workspace + backward_table_offset(n));
    if (n == 0) {
        if (gsigma) hipLaunchKernelGGL(sum_f64_kernel, dim3(1), dim3(256), 0, stream,
                                       (const double*)workspace, 0, 1, gsigma);
        return SMX_OK;
    }
    BwdArgs B;
    B.n = n; B.x = x; B.y = y; B.z = z; B.I = I; B.frame_id = frame_id; B.frames = frames_dev;
    B.nframes = nframes; B.P = *params; B.grad_mu = grad_mu; B.gx = gx; B.gy = gy; B.gz = gz;
    B.gI = gI; B.gsig_part = gsigma ? (double*)workspace : nullptr;
    B.table = table;
    hipLaunchKernelGGL(gauss_table_kernel, dim3((TABLE_ENTRIES + 255) / 256), dim3(256), 0, stream,
                       *params, table);
    SMX_CHECK_LAUNCH("gauss_table_kernel");
    hipLaunchKernelGGL(gauss_backward_kernel, dim3((n + 15) / 16), dim3(256), 0, stream, B);
    SMX_CHECK_LAUNCH("gauss_backward_kernel");
    if (gsigma) {
        // two passes: SUM_BLOCKS partial sums after the n per-emitter terms,
        // then their sum
        double* part = (double*)workspace + n;
        const int per = (n + SUM_BLOCKS - 1) / SUM_BLOCKS;
        hipLaunchKernelGGL(sum_f64_kernel, dim3(SUM_BLOCKS), dim3(256), 0, stream,
                           (const double*)workspace, n, per, part);
        SMX_CHECK_LAUNCH("sum_f64_kernel");
        hipLaunchKernelGGL(sum_f64_kernel, dim3(1), dim3(256), 0, stream, (const double*)part,
                           SUM_BLOCKS, SUM_BLOCKS, gsigma);
        SMX_CHECK_LAUNCH("sum_f64_kernel");
    }
    return SMX_OK;
}

int smx_gauss_backward_binned(int n, const float* x, const float* y, const float* z,
                              const float* I, const int* frame_id, const smx_fov_t* frames_dev,
                              int nframes, int n_tiles, const smx_gauss_params_t* params,
                              long long capacity, const float* grad_mu, float* gx, float* gy,
                              float* gz, float* gI, double* gsigma, void* workspace,
                              size_t workspace_bytes, int* status_dev, hipStream_t stream) {
    if (n < 0 || nframes <= 0 || n_tiles <= 0 || !frames_dev || !grad_mu || !params_ok(params) ||
        capacity < 0 || (n > 0 && (!x || !y || !z || !I)) || !workspace || !status_dev ||
        workspace_bytes < gauss_layout(n, n_tiles, capacity, nullptr, nullptr)) {
        smx::set_error("smx_gauss_backward_binned: invalid arguments (n=%d, nframes=%d, "
                       "n_tiles=%d, workspace %zu bytes)", n, nframes, n_tiles, workspace_bytes);
        return SMX_EVALUE;
    }
    GaussWorkspace ws;
    gauss_layout(n, n_tiles, capacity, (char*)workspace, &ws);
    BwdTileArgs B;
    B.n = n; B.x = x; B.y = y; B.z = z; B.I = I; B.frame_id = frame_id; B.frames = frames_dev;
    B.nframes = nframes; B.n_tiles = n_tiles; B.tile_fov = ws.tile_fov;
    B.tile_begin = ws.tile_begin; B.list = ws.list; B.capacity = capacity; B.P = *params;
    B.table = ws.table; B.hdr = ws.hdr; B.grad_mu = grad_mu;
    B.gx = gx; B.gy = gy; B.gz = gz; B.gI = gI; B.tile_gsig = ws.tile_gsig; B.status = status_dev;
    B.xcd = xcd_map(nframes);
    const int sweep = (int)std::min<long long>(((long long)n + BT_THREADS - 1) / BT_THREADS, 1024LL);
    hipLaunchKernelGGL(gauss_backward_tile_kernel, dim3(n_tiles + std::max(sweep, 1)),
                       dim3(BT_THREADS), 0, stream, B);
    SMX_CHECK_LAUNCH("gauss_backward_tile_kernel");
    if (gsigma) {
        hipLaunchKernelGGL(sum_f64_kernel, dim3(1), dim3(256), 0, stream,
                           (const double*)ws.tile_gsig, n_tiles, n_tiles, gsigma);
        SMX_CHECK_LAUNCH("sum_f64_kernel");
    }
    return SMX_OK;
}

}  // extern "C"
