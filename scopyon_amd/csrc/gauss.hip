// Separable Gaussian image model: forward image, Poisson negative
// log-likelihood with its image gradient, and the hand-written backward with
// respect to emitter positions, depths, intensities and the PSF width.
//
// Reference: the Gaussian radial profile exp(-r^2 / (2 w^2)) / (2 pi w^2) of
// src/scopyon/_epifm.py:128-134 (depth independent), integrated over each
// pixel; the TIRF amplitude exp(-|depth| / penetration_depth) of
// _epifm.py:1289.  Units here are pixels (x, y, sigma) and micrometres (z).
//
// Forward: emitters are binned to 32 x 32 tiles, each tile's list is sorted by
// emitter index, one wave per 16 x 16 quadrant adds every emitter's
// contributions in emitter order (deterministic, no float atomics in global
// memory).  The per-axis pixel factors G(i; x) are evaluated once per emitter
// and tile chunk into LDS (2 x (2R + 1) erf differences); a contribution is
// then one product.
//
// Backward: one 16-lane group per emitter.  With g = dL/dmu over the
// emitter's (2R+1)^2 window, every gradient is a separable sum
//   sum_ij g_ij u_i v_j  (u, v in {G, dG/dx, dG/dsigma}),
// so each lane owns one window column, walks the rows (one coalesced load
// per row), and the group reduces five dot products with shuffles.
#include "common.hpp"
#include "scopyon_hip_types.h"
#include "tiles.hpp"

#include <hipcub/hipcub.hpp>
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
constexpr float INV_SQRT2PI = 0.39894228040143268f;

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
// Gaussian), dG/dx and dG/dsigma.  GQ + 1 phases d = q / GQ are evaluated in
// float64 once per sigma (gauss_table_kernel) and read with linear
// interpolation in d (error ~1e-8 of the peak factor at GQ = 1024 and sigma
// >= 0.5 px: the second derivative in d is bounded by 0.24 / sigma^2).
constexpr int GQ = 1024;
constexpr int GB = 16;            // factor slots per phase (window <= 15)
constexpr int TABLE_ENTRIES = (GQ + 1) * GB;

__global__ void gauss_table_kernel(smx_gauss_params_t P, float4* __restrict__ tab) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= TABLE_ENTRIES) return;
    const int q = idx / GB, b = idx % GB;
    const int R = P.radius, W = 2 * R + 1;
    float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (b < W) {
        const double sigma = (double)P.sigma;
        const double e0 = (double)(b - R) - (double)q / GQ, e1 = e0 + 1.0;
        const double s = 0.70710678118654752 / sigma;
        double g;
        if (e0 >= 0.0) g = 0.5 * (erfc(e0 * s) - erfc(e1 * s));
        else if (e1 <= 0.0) g = 0.5 * (erfc(-e1 * s) - erfc(-e0 * s));
        else g = 0.5 * (erf(e1 * s) - erf(e0 * s));
        const double c = 0.39894228040143268 / sigma;
        const double p0 = c * exp(-0.5 * (e0 / sigma) * (e0 / sigma));
        const double p1 = c * exp(-0.5 * (e1 / sigma) * (e1 / sigma));
        v = make_float4((float)g, (float)(p0 - p1), (float)(-(e1 * p1 - e0 * p0) / sigma), 0.0f);
    }
    tab[idx] = v;
}

// (G, dG/dx, dG/dsigma) of window pixel b at sub-pixel offset d in [0, 1).
__device__ inline float4 table_factor(const float4* __restrict__ tab, float d, int b) {
    const float t = d * (float)GQ;
    const int q = min((int)t, GQ - 1);
    const float f = t - (float)q;
    const float4 a = tab[q * GB + b], c = tab[(q + 1) * GB + b];
    return make_float4(a.x + f * (c.x - a.x), a.y + f * (c.y - a.y), a.z + f * (c.z - a.z), 0.0f);
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

template <int EPT>
__global__ void __launch_bounds__(smx::BIN_THREADS) gauss_bin_count_kernel(
        int n, const float* __restrict__ x, const float* __restrict__ y,
        const int* __restrict__ frame_id, const smx_fov_t* __restrict__ frames, int nframes, int R,
        int n_tiles, int* __restrict__ tile_count, int cs) {
    __shared__ int s_hist[smx::BIN_WIN];
    smx::bin_hist_clear(s_hist);
    const int wbase = gauss_window_base(frame_id, frames, nframes, n, EPT);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const long long i = ((long long)blockIdx.x * EPT + e) * smx::BIN_THREADS + threadIdx.x;
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

struct FwdArgs {
    const float* x;
    const float* y;
    const float* z;
    const float* I;
    const smx_fov_t* frames;
    int nframes;
    const int* tile_fov;
    const int* tile_begin;
    const int* list;
    long long capacity;
    smx_gauss_params_t P;
    float* mu;
    int* status;
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

    const int tile = blockIdx.x;
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
    if (count >= 1 && count <= WAVE_SORT_MAX) {
        if (wave == 0) wave_sort_list(list, count, lane, s_sort);
        __syncthreads();
    } else if (count >= 1 && sorted) {
        int np2 = 1;
        while (np2 < count) np2 <<= 1;
        for (int k = threadIdx.x; k < np2; k += blockDim.x) s_sort[k] = k < count ? list[k] : 0x7fffffff;
        __syncthreads();
        lds_bitonic_sort(s_sort, np2);
    } else if (!sorted && threadIdx.x == 0) {
        atomicOr(A.status, 1);
    }

    for (int base = 0; base < count; base += FCHUNK) {
        const int nc = min(FCHUNK, count - base);
        if (threadIdx.x < nc) {
            const int e = sorted ? s_sort[base + threadIdx.x] : list[base + threadIdx.x];
            const float xe = A.x[e], ye = A.y[e];
            s_xy[threadIdx.x] = make_float2(xe, ye);
            s_lo[threadIdx.x] = make_int2((int)floorf(xe) - R, (int)floorf(ye) - R);
            s_amp[threadIdx.x] = A.I[e] * expf(-fabsf(A.z[e]) * A.P.inv_depth);
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

// Poisson NLL per frame, sum over pixels of mu - k log(mu) (float64), and
// dL/dmu = 1 - k / mu.  Block b of frame f reduces pixels [b * per, (b+1) * per).
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
        if (grad) grad[o] = 1.0f - kk / (m > 0.0f ? m : 1.0e-30f);
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
    const float4* table; // window factors by phase (gauss_table_kernel)
};

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
    float gxr = 0.0f, dgxr = 0.0f, sgxr = 0.0f, gyc = 0.0f, dgyc = 0.0f, sgyc = 0.0f;
    if (valid && j < W) {
        const float4 fx = table_factor(A.table, xe - floorf(xe), j);
        const float4 fy = table_factor(A.table, ye - floorf(ye), j);
        gxr = fx.x;
        dgxr = fx.y;
        sgxr = fx.z;
        gyc = fy.x;
        dgyc = fy.y;
        sgyc = fy.z;
    }
    // the group's row factors go through LDS: one 16-byte broadcast read per
    // row instead of three cross-lane shuffles
    float4* rowf = s_rowf[threadIdx.x >> 4];
    if (j < W) rowf[j] = make_float4(gxr, dgxr, sgxr, 0.0f);
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
        t1 += g * fk.x;
        t2 += g * fk.y;
        t3 += g * fk.z;
    }
    float pI = t1 * gyc, px = t2 * gyc, py = t1 * dgyc, ps = t3 * gyc + t1 * sgyc;
#pragma unroll
    for (int m = 8; m > 0; m >>= 1) {
        pI += __shfl_xor(pI, m);
        px += __shfl_xor(px, m);
        py += __shfl_xor(py, m);
        ps += __shfl_xor(ps, m);
    }
    if (j == 0 && e < A.n) {
        const float a = valid ? expf(-fabsf(ze) * A.P.inv_depth) : 0.0f;
        const float Ia = Ie * a;
        const float sgn = ze > 0.0f ? 1.0f : (ze < 0.0f ? -1.0f : 0.0f);
        if (A.gI) A.gI[e] = a * pI;
        if (A.gx) A.gx[e] = Ia * px;
        if (A.gy) A.gy[e] = Ia * py;
        if (A.gz) A.gz[e] = -sgn * A.P.inv_depth * Ia * pI;
        if (A.gsig_part) A.gsig_part[e] = (double)(Ia * ps);
    }
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
// over the tile's emitters e (zero outside each emitter's window): one
// 32 x 32 MFMA accumulator per tile, emitters streamed 16 at a time in index
// order.  Inputs are rounded to the operand type (float32, float16 or
// bfloat16); accumulation is float32.
// --------------------------------------------------------------------------
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// Per emitter: the window origin and the (2R+1)-pixel factors along each
// axis from the phase table (gauss_table_kernel).  32 lanes per emitter: lane
// t handles window pixel t % 16 of axis t / 16; a block of 256 threads takes
// 8 emitters at a time, striding over the launch.
__global__ void __launch_bounds__(256) sep_factor_kernel(int n, const float* __restrict__ x,
                                                         const float* __restrict__ y,
                                                         const float* __restrict__ z,
                                                         const float* __restrict__ I,
                                                         smx_gauss_params_t P,
                                                         const float4* __restrict__ tab,
                                                         float* __restrict__ fac,
                                                         int2* __restrict__ lo_out) {
    const int t = threadIdx.x & 31, ax = t >> 4, b = t & 15;
    const int R = P.radius, W = 2 * R + 1;
    for (int e = blockIdx.x * 8 + (threadIdx.x >> 5); e < n; e += gridDim.x * 8) {
        const float pos = ax ? y[e] : x[e];
        const bool fin = finite_coord(pos);
        const float fl = floorf(pos);
        float g = 0.0f;
        if (fin && b < W) g = table_factor(tab, pos - fl, b).x;
        float amp = 1.0f;
        if (ax == 0) amp = I[e] * (z ? expf(-fabsf(z[e]) * P.inv_depth) : 1.0f);
        fac[(size_t)e * 32 + t] = ax == 0 ? amp * g : g;
        if (t == 0) {
            const float xv = x[e], yv = y[e];
            lo_out[e] = make_int2(finite_coord(xv) ? (int)floorf(xv) - R : 0,
                                  finite_coord(yv) ? (int)floorf(yv) - R : 0);
        }
    }
}

struct SepArgs {
    const float* fac;
    const int2* lo;
    const smx_fov_t* frames;
    int nframes;
    const int* tile_fov;
    const int* tile_begin;
    const int* list;
    long long capacity;
    int W;
    float background;
    float* mu;           // may be NULL when the NLL is fused (NLL)
    int* status;
    const float* k;      // NLL: observed counts per pixel
    float* grad_mu;      // NLL: dL/dmu = 1 - k / mu (may be NULL)
    double* tile_nll;    // NLL: sum over the tile of mu - k log(mu)
};

template <typename T> struct Frag8 { typedef T type __attribute__((ext_vector_type(8))); };

// DT: operand type, 0: float32 (32x32x2), 1: float16 (32x32x16), 2: bfloat16
// (32x32x16).  NLL: the Poisson NLL of the tile (float64 log, the formula of
// gauss_nll_kernel) and its image gradient in the epilogue, from the
// accumulators, instead of a pass over the written image.
template <int DT, bool NLL>
__global__ void __launch_bounds__(64) sep_gemm_kernel(SepArgs A) {
    __shared__ int s_sort[PSORT_CAP];
    __shared__ float s_fac[16][33];
    __shared__ int2 s_lo[16];
    const int tile = blockIdx.x;
    const int fr = A.tile_fov[tile];
    if (fr < 0 || fr >= A.nframes) return;
    const smx_fov_t F = A.frames[fr];
    const int local = tile - F.tile_off;
    const int tr = local / F.tiles_c, tc = local % F.tiles_c;
    const int l = threadIdx.x;
    const int m = l & 31, half = l >> 5;
    const int row = tr * TS + m, col = tc * TS + m;   // A row / B column of this lane

    const int begin = A.tile_begin[tile];
    int count = A.tile_begin[tile + 1] - begin;
    if ((long long)begin + count > A.capacity) count = (int)max(0LL, A.capacity - begin);
    const int* list = A.list + begin;
    const bool sorted = count <= PSORT_CAP;
    if (count >= 1 && count <= WAVE_SORT_MAX) {
        wave_sort_list(list, count, l, s_sort);
        __syncthreads();
    } else if (count >= 1 && sorted) {
        int np2 = 1;
        while (np2 < count) np2 <<= 1;
        for (int k = l; k < np2; k += 64) s_sort[k] = k < count ? list[k] : 0x7fffffff;
        __syncthreads();
        lds_bitonic_sort(s_sort, np2);
    } else if (!sorted && l == 0) {
        atomicOr(A.status, 1);
    }
    f32x16 acc;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.0f;
    const int W = A.W;
    // factors of chunk `base` (16 emitters x 32 floats, 8 per lane) are
    // fetched into registers one chunk ahead, so their latency overlaps the
    // previous chunk's fragment build and MFMA
    float pre[8];
    int2 pre_lo = make_int2(0, 0);   // lanes 0..15: window origin of emitter `l`
    // branch-free: indices clamped to the list, out-of-chunk values masked,
    // so the eight loads issue back to back
    auto fetch = [&](int base) {
        const int last = count - 1;
        int e[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int k = 2 * r + (l >> 5);
            e[r] = sorted ? s_sort[min(base + k, last)] : list[min(base + k, last)];
        }
        const int el = sorted ? s_sort[min(base + (l & 15), last)] : list[min(base + (l & 15), last)];
#pragma unroll
        for (int r = 0; r < 8; ++r) pre[r] = A.fac[(size_t)e[r] * 32 + (l & 31)];
        pre_lo = A.lo[el];
#pragma unroll
        for (int r = 0; r < 8; ++r)
            if (base + 2 * r + (l >> 5) > last) pre[r] = 0.0f;
    };
    if (count > 0) fetch(0);
    for (int base = 0; base < count; base += 16) {
        const int nk = min(16, count - base);
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int idx = r * 64 + l, k = idx >> 5, t = idx & 31;
            s_fac[k][t] = pre[r];
        }
        if (l < nk) s_lo[l] = pre_lo;
        __syncthreads();
        if (base + 16 < count) fetch(base + 16);
        // operands: window-origin reads first, then clamped factor reads,
        // masked outside the window (no branches between the LDS reads)
        if (DT == 0) {
            int2 o[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) o[j] = s_lo[2 * j + half];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int k = 2 * j + half;
                const int ru = row - o[j].x, cv = col - o[j].y;
                const bool ok = k < nk;
                const float a = s_fac[k][min(max(ru, 0), 15)];
                const float b = s_fac[k][16 + min(max(cv, 0), 15)];
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(
                    (ok && (unsigned)ru < (unsigned)W) ? a : 0.0f,
                    (ok && (unsigned)cv < (unsigned)W) ? b : 0.0f, acc, 0, 0, 0);
            }
        } else {
            typedef typename std::conditional<DT == 1, _Float16, __bf16>::type E;
            typedef typename Frag8<E>::type V8;
            V8 a, b;
            int2 o[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) o[t] = s_lo[half * 8 + t];
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int k = half * 8 + t;
                const int ru = row - o[t].x, cv = col - o[t].y;
                const bool ok = k < nk;
                const float av = s_fac[k][min(max(ru, 0), 15)];
                const float bv = s_fac[k][16 + min(max(cv, 0), 15)];
                a[t] = (E)((ok && (unsigned)ru < (unsigned)W) ? av : 0.0f);
                b[t] = (E)((ok && (unsigned)cv < (unsigned)W) ? bv : 0.0f);
            }
            if constexpr (DT == 1) acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, acc, 0, 0, 0);
            else acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    // acc[i] holds row 8 (i / 4) + 4 half + i % 4, column m of the tile
    const int pj = tc * TS + m;
    double part = 0.0;
    if (pj < F.nh) {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int pi = tr * TS + 8 * (i >> 2) + 4 * half + (i & 3);
            if (pi >= F.nw) continue;
            const size_t o = (size_t)F.pix_off + (size_t)pi * F.nh + pj;
            const float mv = A.background + acc[i];
            if (A.mu) A.mu[o] = mv;
            if (NLL) {
                const float kk = A.k[o];
                const float mc = mv > 0.0f ? mv : 1.0e-30f;
                const double md = (double)mc;
                part += md - (double)kk * log(md);
                if (A.grad_mu) A.grad_mu[o] = 1.0f - kk / mc;
            }
        }
    }
    if (NLL) {
        // fixed reduction tree over the wave: the tile sum is deterministic
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
        if (l == 0) A.tile_nll[tile] = part;
    }
}

// loss[f] = sum of frame f's tile sums in tile order.
__global__ void gauss_tile_nll_finish_kernel(const smx_fov_t* __restrict__ frames, int nframes,
                                             const double* __restrict__ tile_nll,
                                             double* __restrict__ loss) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nframes) return;
    const smx_fov_t F = frames[f];
    double s = 0.0;
    for (int t = 0; t < F.tiles_r * F.tiles_c; ++t) s += tile_nll[F.tile_off + t];
    loss[f] = s;
}

struct GaussWorkspace {
    int* tile_count;   // (n_tiles + 1) x COUNT_STRIDE (padded counters)
    int* tile_dense;   // n_tiles + 1
    int* tile_begin;   // n_tiles + 1
    int* tile_fov;     // n_tiles
    int* list;         // capacity
    double* gsig;      // n
    float* fac;        // n x 32: amp * G along rows (16), G along columns (16)
    int2* lo;          // n: window origin
    float4* table;     // TABLE_ENTRIES window factors by phase
    double* tile_nll;  // n_tiles: fused NLL per tile
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
    size_t o_sig = take(sizeof(double) * (size_t)(n > 0 ? n : 1));
    size_t o_fac = take(sizeof(float) * 32 * (size_t)(n > 0 ? n : 1));
    size_t o_lo = take(sizeof(int2) * (size_t)(n > 0 ? n : 1));
    size_t o_tab = take(sizeof(float4) * (size_t)TABLE_ENTRIES);
    size_t o_tnll = take(sizeof(double) * (size_t)n_tiles);
    size_t cb = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, cb, (int*)nullptr, (int*)nullptr, n_tiles + 1);
    size_t o_cub = take(cb);
    if (ws && base) {
        ws->tile_count = (int*)(base + o_count);
        ws->tile_dense = (int*)(base + o_dense);
        ws->tile_begin = (int*)(base + o_begin);
        ws->tile_fov = (int*)(base + o_fov);
        ws->list = (int*)(base + o_list);
        ws->gsig = (double*)(base + o_sig);
        ws->fac = (float*)(base + o_fac);
        ws->lo = (int2*)(base + o_lo);
        ws->table = (float4*)(base + o_tab);
        ws->tile_nll = (double*)(base + o_tnll);
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
    return backward_table_offset(n) + sizeof(float4) * (size_t)TABLE_ENTRIES;
}

}  // extern "C"

namespace {

// Tile binning shared by both forward forms: count -> scan -> fill.
int gauss_bin(int n, const float* x, const float* y, const int* frame_id,
              const smx_fov_t* frames_dev, int nframes, int n_tiles, int R, long long capacity,
              GaussWorkspace& ws, int* status_dev, hipStream_t stream) {
    // four emitters per thread for large launches (denser LDS windows)
    const int ept = n >= (1 << 19) ? 4 : 1;
    const int nbb = (int)(((long long)n + (long long)smx::BIN_THREADS * ept - 1) /
                          ((long long)smx::BIN_THREADS * ept));
    const int cs = smx::count_stride(2LL * n, n_tiles);
    hipLaunchKernelGGL(smx::tiles_init_kernel, dim3((n_tiles + 1 + 255) / 256), dim3(256), 0,
                       stream, nframes, frames_dev, n_tiles, ws.tile_count, cs, ws.tile_fov);
    SMX_CHECK_LAUNCH("tiles_init_kernel");
    if (n > 0) {
        if (ept == 4)
            hipLaunchKernelGGL(gauss_bin_count_kernel<4>, dim3(nbb), dim3(smx::BIN_THREADS), 0,
                               stream, n, x, y, frame_id, frames_dev, nframes, R, n_tiles,
                               ws.tile_count, cs);
        else
            hipLaunchKernelGGL(gauss_bin_count_kernel<1>, dim3(nbb), dim3(smx::BIN_THREADS), 0,
                               stream, n, x, y, frame_id, frames_dev, nframes, R, n_tiles,
                               ws.tile_count, cs);
        SMX_CHECK_LAUNCH("gauss_bin_count_kernel");
    }
    if (n_tiles + 1 <= smx::SCAN_MAX) {
        // one workgroup scans the counters in place of compact + device scan
        hipLaunchKernelGGL(smx::scan_counts_kernel, dim3(1), dim3(smx::SCAN_THREADS), 0, stream,
                           ws.tile_count, cs, n_tiles + 1, ws.tile_begin);
        SMX_CHECK_LAUNCH("scan_counts_kernel");
    } else {
        size_t cb = ws.cub_bytes;
        hipLaunchKernelGGL(smx::compact_counts_kernel, dim3((n_tiles + 1 + 255) / 256), dim3(256),
                           0, stream, ws.tile_count, cs, n_tiles + 1, ws.tile_dense);
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
    rc = gauss_bin(n, x, y, frame_id, frames_dev, nframes, n_tiles, params->radius, capacity, ws,
                   status_dev, stream);
    if (rc) return rc;
    FwdArgs A;
    A.x = x; A.y = y; A.z = z; A.I = I; A.frames = frames_dev; A.nframes = nframes;
    A.tile_fov = ws.tile_fov; A.tile_begin = ws.tile_begin; A.list = ws.list;
    A.capacity = capacity; A.P = *params; A.mu = mu; A.status = status_dev;
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
    int rc = gauss_bin(n, x, y, frame_id, frames_dev, nframes, n_tiles, params->radius, capacity,
                       ws, status_dev, stream);
    if (rc) return rc;
    if (n > 0) {
        hipLaunchKernelGGL(gauss_table_kernel, dim3((TABLE_ENTRIES + 255) / 256), dim3(256), 0,
                           stream, *params, ws.table);
        SMX_CHECK_LAUNCH("gauss_table_kernel");
        const int blocks = (int)min(((long long)n + 7) / 8, 4096LL);
        hipLaunchKernelGGL(sep_factor_kernel, dim3(blocks), dim3(256), 0, stream, n, x, y, z, I,
                           *params, ws.table, ws.fac, ws.lo);
        SMX_CHECK_LAUNCH("sep_factor_kernel");
    }
    SepArgs A;
    A.fac = ws.fac; A.lo = ws.lo; A.frames = frames_dev; A.nframes = nframes;
    A.tile_fov = ws.tile_fov; A.tile_begin = ws.tile_begin; A.list = ws.list;
    A.capacity = capacity; A.W = 2 * params->radius + 1; A.background = params->background;
    A.mu = mu; A.status = status_dev;
    A.k = k; A.grad_mu = grad_mu; A.tile_nll = ws.tile_nll;
    auto launch = [&](auto nll) {
        constexpr bool L = decltype(nll)::value;
        if (dtype == 0) hipLaunchKernelGGL((sep_gemm_kernel<0, L>), dim3(n_tiles), dim3(64), 0, stream, A);
        else if (dtype == 1) hipLaunchKernelGGL((sep_gemm_kernel<1, L>), dim3(n_tiles), dim3(64), 0, stream, A);
        else hipLaunchKernelGGL((sep_gemm_kernel<2, L>), dim3(n_tiles), dim3(64), 0, stream, A);
    };
    if (k) launch(std::true_type());
    else launch(std::false_type());
    SMX_CHECK_LAUNCH("sep_gemm_kernel");
    if (k) {
        hipLaunchKernelGGL(gauss_tile_nll_finish_kernel, dim3((nframes + 63) / 64), dim3(64), 0,
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
    float4* table = (float4*)((char*)workspace + backward_table_offset(n));
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

}  // extern "C"
