// Emitter weighting and pixel integration ("splat") for one or many fields
// of view (FOVs).
//
// Reference: src/scopyon/_epifm.py:1267-1319 (per-emitter weight),
//            _epifm.py:215-282 (overlay_signal_: per pixel, the sum of the
//            1 nm cartesian PSF cells inside the pixel).
//
// Pixel sums come from a float64 summed-area table (SAT) per depth key:
// sum over cells [B0,B1) x [C0,C1) = S[B1][C1] - S[B0][C1] - S[B1][C0] + S[B0][C0].
// The cell boundaries B, C are computed with the reference's float64
// expression ceil((k*pl - o)/res) and the same end clamps.
//
// Work decomposition: the frame is cut into 32x32-pixel tiles; emitters are
// binned to the tiles their footprint touches (count -> scan -> fill), each
// tile's list is sorted by emitter index in LDS (so every pixel adds its
// contributions in emitter order, like the reference loop, and the result is
// run-to-run deterministic), then one 256-thread workgroup per tile
// accumulates its 1024 pixels in registers.  No float atomics.
#include <cstdlib>
#include <mutex>
#include <type_traits>
#include "common.hpp"
#include "emccd_block.hpp"
#include "scopyon_hip.h"
#include "scopyon_hip_types.h"
#include "tiles.hpp"

#include <hipcub/hipcub.hpp>

namespace {

using smx::align256;
using smx::fill_int_kernel;
using smx::lds_add;
using smx::lds_bitonic_sort;
using smx::wave_sort_list;
using smx::WAVE_SORT_MAX;
using smx::tile_fov_kernel;

constexpr int TS = 32;         // tile side in pixels
constexpr int TB = TS + 1;     // boundaries per tile axis
constexpr int PY = 4;          // pixel rows per thread (256 threads x 4 = 32 x 32)
constexpr int RS = TS / PY;    // row stride between a thread's pixels
constexpr int CHUNK = 4;       // emitters staged in LDS per round (general path)
constexpr int SORT_CAP = 4096; // max tile list sorted in LDS

// One axis of the footprint: pixel range [k0, k1) and its cell boundaries.
struct Axis {
    double o;      // (half + xi) - ext_half : object coordinate of table cell 0
    long long k0, k1;
};

__device__ inline Axis make_axis(double half, double xi, double ext_half, double pl, int npix) {
    Axis a;
    double base = half + xi;
    a.o = base - ext_half;
    long long lo = (long long)floor(a.o / pl);
    long long hi = (long long)ceil((base + ext_half) / pl);
    a.k0 = lo > 0 ? lo : 0;
    a.k1 = hi < npix ? hi : npix;
    return a;
}

// Boundary for pixel edge k (k0 <= k <= k1), with the reference's clamps.
__device__ inline int axis_boundary(const Axis& a, long long k, double pl, double res, int ntab) {
    long long b = smx::ceil_i64(((double)k * pl - a.o) / res);
    if (k == a.k0 && b < 0) b = 0;
    if (k == a.k1 && b > ntab) b = ntab;
    return (int)b;
}

// --------------------------------------------------------------------------
// Emitter weights: _epifm.py:1280-1315 (+ bleaching budget, :1294-1306)
// --------------------------------------------------------------------------
__global__ void weights_kernel(int n, const double* __restrict__ depth,
                               const double* __restrict__ p_state, smx_weight_params_t P,
                               double unit_time, double* __restrict__ budget,
                               const int* __restrict__ budget_slot, const int* __restrict__ slot_map,
                               int n_slot_map, double* __restrict__ w_out,
                               int* __restrict__ key_out, double* __restrict__ pstate_out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double d = depth[i] - P.focal_depth;
    double amp = P.amplitude * exp(-fabs(d) / P.penetration_depth);
    double n_abs = amp * P.x_sec * unit_time;
    double n_emit = P.quantum_yield * n_abs * P.one_minus_10_pow_minus_a;
    double ps = p_state ? p_state[i] : 1.0;
    if (budget) {
        int s = budget_slot ? budget_slot[i] : i;
        double b = budget[s] - n_emit;
        if (b <= 0.0) {
            b = 0.0;
            ps = 0.0;
        }
        budget[s] = b;
    }
    w_out[i] = P.spectral_psf_norm * ((ps * n_emit) / P.four_pi);
    if (pstate_out) pstate_out[i] = ps;
    // depth key (_epifm.py:76-85): int(|d| / res) below cutoff + res, else the
    // cutoff key; then the table slot holding that key
    double ad = fabs(d);
    long long key = (ad < P.depth_cutoff + P.res) ? (long long)(ad / P.res) : P.cutoff_slot;
    if (slot_map) {
        if (key < 0) key = 0;
        if (key >= n_slot_map) key = n_slot_map - 1;
        key = slot_map[key];
    }
    key_out[i] = (int)key;
}

// --------------------------------------------------------------------------
// Binning
// --------------------------------------------------------------------------
struct TileRange { int r0, r1, c0, c1; };   // inclusive tile index ranges

__device__ inline bool emitter_tiles(const smx_fov_t& F, double xi, double yi, double ext_half,
                                     TileRange& t) {
    Axis ax = make_axis(F.half_w, xi, ext_half, F.pixel_length, F.nw);
    Axis ay = make_axis(F.half_h, yi, ext_half, F.pixel_length, F.nh);
    if (ax.k1 <= ax.k0 || ay.k1 <= ay.k0) return false;
    t.r0 = (int)(ax.k0 / TS);
    t.r1 = (int)((ax.k1 - 1) / TS);
    t.c0 = (int)(ay.k0 / TS);
    t.c1 = (int)((ay.k1 - 1) / TS);
    return true;
}

__device__ inline int fov_of(const int* __restrict__ fov_id, int i) { return fov_id ? fov_id[i] : 0; }
__device__ inline bool fov_ok(int f, int nfov) { return f >= 0 && f < nfov; }

__global__ void bin_count_kernel(int n, const double* __restrict__ u, const double* __restrict__ v,
                                 const double* __restrict__ w, const int* __restrict__ fov_id,
                                 const smx_fov_t* __restrict__ fovs, int nfov, double ext_half,
                                 int n_tiles, int* __restrict__ tile_count, int cs) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (!(w[i] > 0.0)) return;
    const int f = fov_of(fov_id, i);
    if (!fov_ok(f, nfov)) return;
    const smx_fov_t F = fovs[f];
    TileRange t;
    if (!emitter_tiles(F, u[i] - F.p0u, v[i] - F.p0v, ext_half, t)) return;
    for (int r = t.r0; r <= t.r1; ++r)
        for (int c = t.c0; c <= t.c1; ++c)
        {
            int tile = F.tile_off + r * F.tiles_c + c;
            if (tile >= 0 && tile < n_tiles) atomicAdd(&tile_count[(size_t)tile * cs], 1);
        }
}

__global__ void bin_fill_kernel(int n, const double* __restrict__ u, const double* __restrict__ v,
                                const double* __restrict__ w, const int* __restrict__ fov_id,
                                const smx_fov_t* __restrict__ fovs, int nfov, double ext_half,
                                int n_tiles, const int* __restrict__ tile_begin,
                                int* __restrict__ tile_left, int cs,
                                int* __restrict__ list, long long capacity, int* __restrict__ status) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (!(w[i] > 0.0)) return;
    const int f = fov_of(fov_id, i);
    if (!fov_ok(f, nfov)) return;
    const smx_fov_t F = fovs[f];
    TileRange t;
    if (!emitter_tiles(F, u[i] - F.p0u, v[i] - F.p0v, ext_half, t)) return;
    for (int r = t.r0; r <= t.r1; ++r)
        for (int c = t.c0; c <= t.c1; ++c) {
            int tile = F.tile_off + r * F.tiles_c + c;
            if (tile < 0 || tile >= n_tiles) continue;
            int pos = atomicSub(&tile_left[(size_t)tile * cs], 1) - 1;   // counts run down to 0
            long long at = (long long)tile_begin[tile] + pos;
            if (at < capacity) list[at] = i;
            else atomicOr(status, 2);
        }
}

// --------------------------------------------------------------------------
// Tile accumulation
// --------------------------------------------------------------------------
struct TileArgs {
    const double* u;
    const double* v;
    const double* w;
    const int* key;          // SAT slot per emitter
    const int* fov_id;
    const smx_fov_t* fovs;
    int nfov;
    const int* tile_fov;     // FOV of each global tile
    const int* tile_begin;   // n_tiles + 1 offsets into list
    int* list;               // sorted in place per tile
    long long capacity;      // entries in list
    const double* sat;
    size_t sat_stride;
    int n_slots;             // tables available in sat
    smx::SatLayout L;
    double res, ext_half, unit_area;
    int ntab;
    int accumulate;          // 1: out += acc ; 0: out = acc
    int xcd;                 // 1: workgroups map to tiles by smx::xcd_tile
    double* out;
    int* status;             // status word (bits as smx_splat_expected documents)
};

__global__ void __launch_bounds__(256) splat_tile_kernel(TileArgs A) {
    // LDS: the sort buffer is only live before the main loop; the corner
    // staging area reuses it afterwards.
    __shared__ union {
        int list[SORT_CAP];
        double corner[CHUNK][TB][TB];
    } u;
    __shared__ int s_id[CHUNK];
    __shared__ int s_k0[CHUNK][2], s_k1[CHUNK][2];
    __shared__ double s_o[CHUNK][2];
    __shared__ int s_B[CHUNK][2][TB];          // cell boundary per tile edge (-1: outside)
    __shared__ long long s_P[CHUNK][2][TB];    // SAT index part of that boundary
    __shared__ double s_w[CHUNK];
    __shared__ long long s_sat_off[CHUNK];
    __shared__ int s_lo[CHUNK][2], s_n[CHUNK][2];   // valid tile-edge range per axis
    __shared__ int s_off[CHUNK + 1];                 // prefix of valid corners per emitter

    const int tile = A.xcd ? smx::xcd_tile(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    const int fov = A.tile_fov[tile];
    if (fov < 0 || fov >= A.nfov) return;   // tile not covered by any FOV
    const smx_fov_t F = A.fovs[fov];
    const int local = tile - F.tile_off;
    const int tr = local / F.tiles_c, tc = local % F.tiles_c;
    const int ty = threadIdx.x / TS, tx = threadIdx.x % TS;   // rows ty + RS*q, column tx
    const int pj = tc * TS + tx;

    const int begin = A.tile_begin[tile];
    int count = A.tile_begin[tile + 1] - begin;
    if ((long long)begin + count > A.capacity) count = (int)max(0LL, A.capacity - begin);
    int* list = A.list + begin;

    // sort the tile list by emitter index (emitter order per pixel) and write
    // it back in place
    if (count > 1 && count <= SORT_CAP) {
        int np2 = 1;
        while (np2 < count) np2 <<= 1;
        for (int k = threadIdx.x; k < np2; k += blockDim.x) u.list[k] = k < count ? list[k] : 0x7fffffff;
        __syncthreads();
        lds_bitonic_sort(u.list, np2);
        for (int k = threadIdx.x; k < count; k += blockDim.x) list[k] = u.list[k];
        __syncthreads();
    }

    double acc[PY];
#pragma unroll
    for (int q = 0; q < PY; ++q) acc[q] = 0.0;
    for (int base = 0; base < count; base += CHUNK) {
        const int nc = min(CHUNK, count - base);
        // phase A0: per emitter and axis, the footprint range and cell origin
        if (threadIdx.x < nc * 2) {
            int c = threadIdx.x >> 1, ax = threadIdx.x & 1;
            int e = list[base + c];
            double xi = ax == 0 ? A.u[e] - F.p0u : A.v[e] - F.p0v;
            Axis a = make_axis(ax == 0 ? F.half_w : F.half_h, xi, A.ext_half, F.pixel_length,
                               ax == 0 ? F.nw : F.nh);
            s_o[c][ax] = a.o;
            s_k0[c][ax] = (int)a.k0;
            s_k1[c][ax] = (int)a.k1;
            long long e0 = (long long)(ax == 0 ? tr : tc) * TS;   // first edge of the tile
            long long lo = max(a.k0 - e0, 0LL), hi = min(a.k1 - e0, (long long)TS);
            s_lo[c][ax] = (int)lo;
            s_n[c][ax] = hi >= lo ? (int)(hi - lo + 1) : 0;
            if (ax == 0) {
                s_id[c] = e;
                int slot = A.key[e];
                bool ok = slot >= 0 && slot < A.n_slots;
                s_w[c] = ok ? A.w[e] : 0.0;
                s_sat_off[c] = ok ? (long long)slot * (long long)A.sat_stride : -1;
                if (!ok) atomicOr(A.status, 8);
            }
        }
        __syncthreads();
        // phase A1: cell boundaries of this tile's edges inside each footprint
        for (int t = threadIdx.x; t < nc * 2 * TB; t += blockDim.x) {
            int c = t / (2 * TB), r = t % (2 * TB);
            int ax = r / TB, m = r % TB;
            long long k = (long long)(ax == 0 ? tr : tc) * TS + m;
            int b = -1;
            long long part = 0;
            if (k >= s_k0[c][ax] && k <= s_k1[c][ax]) {
                Axis a;
                a.o = s_o[c][ax];
                a.k0 = s_k0[c][ax];
                a.k1 = s_k1[c][ax];
                b = axis_boundary(a, k, F.pixel_length, A.res, A.ntab);
                part = (long long)(ax == 0 ? A.L.row_part(b) : A.L.col_part(b));
            }
            s_B[c][ax][m] = b;
            s_P[c][ax][m] = part;
        }
        __syncthreads();
        // phase B: gather only the valid SAT corners (compacted over the chunk),
        // four independent loads in flight per thread
        if (threadIdx.x == 0) {
            int run = 0;
            for (int c = 0; c < nc; ++c) {
                s_off[c] = run;
                run += s_sat_off[c] >= 0 ? s_n[c][0] * s_n[c][1] : 0;
            }
            s_off[nc] = run;
        }
        __syncthreads();
        {
            const int total = s_off[nc];
            for (int t0 = threadIdx.x; t0 < total; t0 += 4 * 256) {
                double vals[4];
                int dst[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    int t = t0 + q * 256;
                    dst[q] = -1;
                    vals[q] = 0.0;
                    if (t < total) {
                        int c = 0;
                        while (c + 1 < nc && s_off[c + 1] <= t) ++c;
                        int r = t - s_off[c];
                        int m = s_lo[c][0] + r / s_n[c][1];
                        int l = s_lo[c][1] + r % s_n[c][1];
                        vals[q] = smx::gload(A.sat + s_sat_off[c] + s_P[c][0][m] + s_P[c][1][l]);
                        dst[q] = (c * TB + m) * TB + l;
                    }
                }
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (dst[q] >= 0) (&u.corner[0][0][0])[dst[q]] = vals[q];
            }
        }
        __syncthreads();
        // phase C: per-pixel accumulation in emitter order
        if (pj < F.nh) {
            for (int c = 0; c < nc; ++c) {
                if (s_sat_off[c] < 0) continue;
                int c0 = s_B[c][1][tx], c1 = s_B[c][1][tx + 1];
                if (c0 < 0 || c1 < 0 || c1 <= c0) continue;
#pragma unroll
                for (int q = 0; q < PY; ++q) {
                    int r = ty + RS * q;
                    int b0 = s_B[c][0][r], b1 = s_B[c][0][r + 1];
                    if (b0 < 0 || b1 < 0 || b1 <= b0) continue;
                    double sv = ((u.corner[c][r + 1][tx + 1] - u.corner[c][r][tx + 1])
                                 - u.corner[c][r + 1][tx]) + u.corner[c][r][tx];
                    double photons = sv * A.unit_area;
                    if (photons > 0.0) acc[q] += photons * s_w[c];
                }
            }
        }
        __syncthreads();
    }
    if (pj < F.nh) {
#pragma unroll
        for (int q = 0; q < PY; ++q) {
            int pi = tr * TS + ty + RS * q;
            if (pi >= F.nw) continue;
            size_t o = (size_t)F.pix_off + (size_t)pi * F.nh + pj;
            if (A.accumulate) {
                if (count > 0) A.out[o] = A.out[o] + acc[q];
            } else {
                A.out[o] = acc[q];
            }
        }
    }
}

// --------------------------------------------------------------------------
// Phase-table fast path (pixel length = P table cells, P integer)
//
// With pl = P * res every emitter's interior cell boundaries along an axis
// are phi + P * (m - 1) for m >= 1 (phi in [1, P]) and 0 for m = 0, clamped
// to the table end -- the reference's ceil((k pl - o)/res) with its two
// clamps.  The pixel integrals therefore depend only on (phi_u, phi_v) and
// the pixel's offset (du, dv) inside the footprint, so they are tabulated
// once per depth key: K[phi_u][phi_v][du][dv] = the same four-corner SAT
// difference the general path evaluates (bit-identical values).  The splat
// then costs one coalesced load per (emitter, pixel).  An emitter whose
// boundaries are not on the regular grid (a float64 rounding tie in
// ceil((k pl - o)/res), probability ~1e-12 per boundary) is flagged and
// integrated through the SAT instead.
// --------------------------------------------------------------------------
struct EmitterRec {
    int lo_u, lo_v;          // first footprint pixel per axis (may be < 0)
    short n_u, n_v;          // footprint pixels per axis (hi - lo)
    short su, sv;            // table element strides per footprint pixel along u, v
    int flags;               // 1: irregular boundaries -> SAT path; 2: no table slot
    unsigned frac;           // LUT mode: sub-phase fractions along u (low 16 bits) and v
    double w;
    long long table_off;     // element of pixel (du, dv) = table_off + du * su + dv * sv
};

// Tiles (side 2^lts) touched by a record's footprint rectangle [lo, lo + n)
// clipped to the frame.
__device__ inline bool rec_tiles(const smx_fov_t& F, const EmitterRec& r, TileRange& t,
                                 int lts = 5) {
    const int r0 = max(r.lo_u, 0), c0 = max(r.lo_v, 0);
    const int r1 = (int)min((long long)r.lo_u + r.n_u, (long long)F.nw) - 1;
    const int c1 = (int)min((long long)r.lo_v + r.n_v, (long long)F.nh) - 1;
    if (r1 < r0 || c1 < c0) return false;
    t.r0 = r0 >> lts; t.r1 = r1 >> lts; t.c0 = c0 >> lts; t.c1 = c1 >> lts;
    return true;
}

__device__ inline int regular_phase(const Axis& a, double half, double xi, double ext_half,
                                    double pl, double res, double inv_res, int ntab, int P, int M,
                                    long long lo, long long hi, bool& regular) {
    // phi = ceil(((lo + 1) pl - o) / res); check every boundary lo+2..hi (the
    // boundary lo+1 is phi itself, also under the clamp at hi)
    const double q1 = (double)(lo + 1) * pl - a.o;
    long long phi = smx::ceil_div(q1, res, inv_res);
    if (phi < 1 || phi > P || hi - lo > M) regular = false;
    // Shortcut, exact: with pl / res = P to 1e-13, boundary k's quotient is
    // q1 / res + P (k - lo - 1) to within 3e-10 (k - lo <= M <= 32, quotients
    // below 1e6), so a first quotient at least 1e-8 from an integer fixes
    // every ceiling to phi + P (k - lo - 1): the loop below would find all of
    // them regular.  Otherwise each boundary is checked.
    const double qq = q1 * inv_res, fr = qq - floor(qq);
    if (regular && M <= 32 && fabs(qq) < 1e6 && fr > 1e-8 && fr < 1.0 - 1e-8 &&
        fabs(pl * inv_res - (double)P) < 1e-13 * (double)P)
        return (int)(phi - 1);
    for (long long k = lo + 2; regular && k <= hi; ++k) {
        long long b = smx::ceil_div((double)k * pl - a.o, res, inv_res);
        long long want = phi + (long long)P * (k - lo - 1);
        if (k == hi) {
            b = b > ntab ? ntab : b;
            want = want > ntab ? ntab : want;
        }
        if (b != want) regular = false;
    }
    return (int)(phi - 1);
}

// Record of emitter i in FOV f (flags 2: no table slot or FOV, not binned).
__device__ inline EmitterRec make_rec(int i, int f, const double* __restrict__ u,
                                      const double* __restrict__ v, const double* __restrict__ w,
                                      const int* __restrict__ key, const smx_fov_t* __restrict__ fovs,
                                      int nfov, double ext_half, double res, int ntab, int P, int M,
                                      int n_slots, size_t k_stride, int folded,
                                      const smx::PhaseFold& PF, int* __restrict__ status) {
    EmitterRec r;
    r.lo_u = r.lo_v = 0;
    r.n_u = r.n_v = 0;
    r.su = (short)M;
    r.sv = 1;
    r.flags = 0;
    r.frac = 0;
    r.w = w[i];
    r.table_off = 0;
    int slot = key[i];
    if (!fov_ok(f, nfov) || slot < 0 || slot >= n_slots) {
        r.flags = 2;
        if (fov_ok(f, nfov) && r.w > 0.0) atomicOr(status, 8);
        return r;
    }
    const smx_fov_t& F = fovs[f];
    const double xi = u[i] - F.p0u, yi = v[i] - F.p0v;
    const double inv_res = 1.0 / res, inv_pl = 1.0 / F.pixel_length;
    bool regular = true;
    int ph[2];
    for (int ax = 0; ax < 2; ++ax) {
        double half = ax == 0 ? F.half_w : F.half_h;
        double x = ax == 0 ? xi : yi;
        double base = half + x;
        Axis a;
        a.o = base - ext_half;
        long long lo = smx::floor_div(a.o, F.pixel_length, inv_pl);
        long long hi = smx::ceil_div(base + ext_half, F.pixel_length, inv_pl);
        ph[ax] = regular_phase(a, half, x, ext_half, F.pixel_length, res, inv_res, ntab, P, M, lo,
                               hi, regular);
        if (lo < -(1LL << 30) || lo > (1LL << 30)) regular = false;
        long long span = hi - lo;
        short len = (short)(span < 0 ? 0 : (span > 32767 ? 32767 : span));
        if (ax == 0) { r.lo_u = (int)lo; r.n_u = len; } else { r.lo_v = (int)lo; r.n_v = len; }
    }
    if (!regular) {
        r.flags = 1;
    } else if (!folded) {
        r.table_off = (long long)slot * (long long)k_stride +
                      ((long long)ph[0] * P + ph[1]) * (long long)(M * M);
    } else {
        // pixels past np(ph) are empty (zero integral): drop them, then address
        // the canonical block, reversed along a reflected axis, transposed when
        // the canonical pair is stored as (v, u)
        const int mf = PF.mf;
        bool ru, rv;
        int cu = PF.canon(ph[0], ru), cv = PF.canon(ph[1], rv);
        int npu = PF.npix(ph[0]), npv = PF.npix(ph[1]);
        if (r.n_u > npu) r.n_u = (short)npu;
        if (r.n_v > npv) r.n_v = (short)npv;
        int a0 = ru ? npu - 1 : 0, sa = ru ? -1 : 1;   // index along u in the block: a0 + sa du
        int b0 = rv ? npv - 1 : 0, sb = rv ? -1 : 1;
        long long off;
        if (cu <= cv) {
            off = (long long)PF.block(cu, cv) * mf * mf + (long long)a0 * mf + b0;
            r.su = (short)(sa * mf);
            r.sv = (short)sb;
        } else {
            off = (long long)PF.block(cv, cu) * mf * mf + (long long)b0 * mf + a0;
            r.su = (short)sa;
            r.sv = (short)(sb * mf);
        }
        r.table_off = (long long)slot * (long long)k_stride + off;
    }
    return r;
}

// Tile counts and slot claims aggregate per block in LDS (BIN_THREADS
// emitters, a window of BIN_WIN tiles from the first tile of the block's
// first FOV): device-scope atomics are the cost of binning.
constexpr int REC_THREADS = smx::BIN_THREADS;
constexpr int REC_WIN = smx::BIN_WIN;

__device__ inline int window_base(const int* __restrict__ fov_id, const smx_fov_t* __restrict__ fovs,
                                  int nfov, int n) {
    const int i0 = blockIdx.x * REC_THREADS;
    const int f0 = i0 < n ? fov_of(fov_id, i0) : -1;
    return fov_ok(f0, nfov) ? fovs[f0].tile_off : 0;
}

// LUT-mode geometry of the records (q = 0: phase tables of the 1 nm path).
struct LutRec {
    int q, m, c;
    size_t slice_elems;
};

// LUT-mode record of emitter i in FOV f: position a = (W/2 + x) / pl in
// pixels from the frame edge (pixel k spans [k, k + 1), _epifm.py:230-231),
// a q = q_i + frac: phase q_i mod q of pixel floor(q_i / q), the footprint
// starts c pixels before it; the table block of phases (q_u, q_v) and the
// 16-bit fractions for the bilinear taps towards (q_u + 1, q_v + 1).
__device__ inline EmitterRec make_rec_lut(int i, int f, const double* __restrict__ u,
                                          const double* __restrict__ v,
                                          const double* __restrict__ w, const int* __restrict__ key,
                                          const smx_fov_t* __restrict__ fovs, int nfov, int n_slots,
                                          const LutRec& LR, int* __restrict__ status) {
    EmitterRec r;
    r.lo_u = r.lo_v = 0;
    r.n_u = r.n_v = 0;
    r.su = (short)(LR.m * 4);   // cell form: 4 floats per pixel
    r.sv = 4;
    r.flags = 0;
    r.frac = 0;
    r.w = w[i];
    r.table_off = 0;
    const int slot = key[i];
    if (!fov_ok(f, nfov) || slot < 0 || slot >= n_slots) {
        r.flags = 2;
        if (fov_ok(f, nfov) && r.w > 0.0) atomicOr(status, 8);
        return r;
    }
    const smx_fov_t& F = fovs[f];
    int ph[2];
    unsigned fr[2];
    for (int ax = 0; ax < 2; ++ax) {
        const double half = ax == 0 ? F.half_w : F.half_h;
        const double x = ax == 0 ? u[i] - F.p0u : v[i] - F.p0v;
        const double t = (half + x) / F.pixel_length * (double)LR.q;
        ph[ax] = 0;
        fr[ax] = 0;
        if (!(fabs(t) < 1.0e9)) continue;        // far off the frame (or NaN): nothing binned
        const double qi = floor(t);
        const long long qq = (long long)qi;
        long long pix = qq / LR.q;
        if (qq - pix * LR.q < 0) --pix;          // floor division
        ph[ax] = (int)(qq - pix * LR.q);
        fr[ax] = (unsigned)((t - qi) * 65535.0 + 0.5);
        const int lo = (int)(pix - LR.c);
        if (ax == 0) { r.lo_u = lo; r.n_u = (short)LR.m; } else { r.lo_v = lo; r.n_v = (short)LR.m; }
    }
    r.frac = fr[0] | (fr[1] << 16);
    r.table_off = (long long)slot * (long long)LR.slice_elems +
                  ((long long)ph[0] * LR.q + ph[1]) * (long long)(LR.m * LR.m * 4);
    return r;
}

__global__ void __launch_bounds__(REC_THREADS) emitter_rec_kernel(
        int n, const double* __restrict__ u, const double* __restrict__ v,
        const double* __restrict__ w, const int* __restrict__ key, const int* __restrict__ fov_id,
        const smx_fov_t* __restrict__ fovs, int nfov, double ext_half, double res, int ntab, int P,
        int M, int n_slots, size_t k_stride, int folded, smx::PhaseFold PF,
        EmitterRec* __restrict__ rec, int n_tiles, int* __restrict__ tile_count, int cs,
        int* __restrict__ status, int lts, LutRec LR) {
    __shared__ int s_hist[REC_WIN];
    smx::bin_hist_clear(s_hist);
    const int wbase = window_base(fov_id, fovs, nfov, n);
    __syncthreads();
    const int i = blockIdx.x * REC_THREADS + threadIdx.x;
    if (i < n) {
        const int f = fov_of(fov_id, i);
        const EmitterRec r =
                LR.q > 0 ? make_rec_lut(i, f, u, v, w, key, fovs, nfov, n_slots, LR, status)
                         : make_rec(i, f, u, v, w, key, fovs, nfov, ext_half, res, ntab, P, M,
                                    n_slots, k_stride, folded, PF, status);
        rec[i] = r;
        TileRange t;
        if (tile_count && r.flags != 2 && r.w > 0.0 && rec_tiles(fovs[f], r, t, lts)) {
            const smx_fov_t& F = fovs[f];
            for (int a = t.r0; a <= t.r1; ++a)
                for (int b = t.c0; b <= t.c1; ++b) {
                    const int tile = F.tile_off + a * F.tiles_c + b;
                    if (tile >= 0 && tile < n_tiles)
                        smx::bin_count(s_hist, wbase, tile, tile_count, cs);
                }
        }
    }
    __syncthreads();
    if (tile_count) smx::bin_count_flush(s_hist, wbase, tile_count, cs);
}

// Tile lists from the records (phase path): the footprint is the record's
// pixel rectangle, already computed with the reference's float64 boundaries.
// Entries of a footprint within 2 x 2 tiles of the window take an LDS rank,
// the block claims each window tile's slots with one global atomic, and the
// entries are written at claim + rank; other entries claim a slot each.
__global__ void __launch_bounds__(REC_THREADS) bin_fill_rec_kernel(
        int n, const EmitterRec* __restrict__ rec, const int* __restrict__ fov_id,
        const smx_fov_t* __restrict__ fovs, int nfov, int n_tiles,
        const int* __restrict__ tile_begin, int* __restrict__ tile_left, int cs,
        int* __restrict__ list, long long capacity, int* __restrict__ status, int lts) {
    __shared__ int s_hist[REC_WIN];
    smx::bin_hist_clear(s_hist);
    const int wbase = window_base(fov_id, fovs, nfov, n);
    __syncthreads();
    const int i = blockIdx.x * REC_THREADS + threadIdx.x;
    int ent[4];   // window entries: tile offset << 16 | LDS rank (-1: none)
#pragma unroll
    for (int q = 0; q < 4; ++q) ent[q] = -1;
    const int f = i < n ? fov_of(fov_id, i) : -1;
    if (fov_ok(f, nfov)) {
        const EmitterRec r = rec[i];
        const smx_fov_t& F = fovs[f];
        TileRange t;
        if (r.flags != 2 && r.w > 0.0 && rec_tiles(F, r, t, lts)) {
            const bool small = t.r1 - t.r0 <= 1 && t.c1 - t.c0 <= 1;
            int q = 0;
            for (int a = t.r0; a <= t.r1; ++a)
                for (int b = t.c0; b <= t.c1; ++b) {
                    const int tile = F.tile_off + a * F.tiles_c + b;
                    if (tile < 0 || tile >= n_tiles) continue;
                    const unsigned d = (unsigned)(tile - wbase);
                    if (small && d < (unsigned)REC_WIN) {
                        // ranks stay below 2^16: at most REC_THREADS entries per tile
                        ent[q & 3] = (int)(d << 16) | atomicAdd(&s_hist[d], 1);
                        ++q;
                    } else {
                        const int pos = atomicSub(&tile_left[(size_t)tile * cs], 1) - 1;
                        const long long at = (long long)tile_begin[tile] + pos;
                        if (at < capacity) list[at] = i;
                        else atomicOr(status, 2);
                    }
                }
        }
    }
    __syncthreads();
    smx::bin_claim(s_hist, wbase, tile_left, cs);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (ent[q] < 0) continue;
        const int d = ent[q] >> 16, rank = ent[q] & 0xffff;
        const long long at = (long long)tile_begin[wbase + d] + s_hist[d] + rank;
        if (at < capacity) list[at] = i;
        else atomicOr(status, 2);
    }
}

// K[slot][phi_u][phi_v][du][dv] from the SATs (same arithmetic as the tile kernel).
__global__ void phase_table_kernel(const double* __restrict__ sat, size_t sat_stride,
                                   smx::SatLayout L, int ntab, int P, int M, int nslots,
                                   double unit_area, double* __restrict__ K) {
    size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t per = (size_t)P * P * M * M;
    if (idx >= per * nslots) return;
    size_t slot = idx / per, r = idx % per;
    int dv = (int)(r % M); r /= M;
    int du = (int)(r % M); r /= M;
    int pv = (int)(r % P), pu = (int)(r / P);
    auto edge = [&](int phi, int m) -> int {
        if (m == 0) return 0;
        long long b = (long long)phi + (long long)P * (m - 1);
        return (int)(b > ntab ? ntab : b);
    };
    int b0 = edge(pu + 1, du), b1 = edge(pu + 1, du + 1);
    int c0 = edge(pv + 1, dv), c1 = edge(pv + 1, dv + 1);
    double val = 0.0;
    if (b1 > b0 && c1 > c0) {
        const double* S = sat + slot * sat_stride;
        val = ((S[L.index(b1, c1)] - S[L.index(b0, c1)]) - S[L.index(b1, c0)]) + S[L.index(b0, c0)];
    }
    // photons = sum * unit_area, kept only when positive (the reference's test)
    double photons = val * unit_area;
    K[idx] = photons > 0.0 ? photons : 0.0;
}

// Folded float32 copy: one mf x mf block per unordered pair of canonical
// phases (smx::PhaseFold), taken from K[slot][ph_u][ph_v][a][b].
__global__ void phase_fold_kernel(const double* __restrict__ K, int nslots, int P, int M,
                                  smx::PhaseFold PF, float* __restrict__ Kf) {
    size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int nc = PF.nc, mf = PF.mf;
    size_t per = (size_t)nc * nc * mf * mf;
    if (idx >= per * nslots) return;
    size_t slot = idx / per, r = idx % per;
    int b = (int)(r % mf); r /= mf;
    int a = (int)(r % mf); r /= mf;
    int cv = (int)(r % nc), cu = (int)(r / nc);
    if (cu > cv) return;
    int pu = PF.uncanon(cu), pv = PF.uncanon(cv);
    size_t src = (size_t)slot * P * P * M * M + (((size_t)pu * P + pv) * M + a) * M + b;
    size_t dst = (size_t)slot * PF.elems() + PF.block(cu, cv) * mf * mf + (size_t)a * mf + b;
    Kf[dst] = (float)K[src];
}

struct FastArgs {
    const void* Kraw;        // phase table (double or float)
    const double* u;
    const double* v;
    const int* key;
    const smx_fov_t* fovs;
    int nfov;
    const int* tile_fov;
    const int* tile_begin;
    int* list;
    long long capacity;
    const EmitterRec* rec;
    const double* K;
    int M;
    const double* sat;
    size_t sat_stride;
    smx::SatLayout L;
    double res, ext_half, unit_area;
    int ntab;
    int accumulate;
    int xcd;                 // 1: workgroups map to tiles by smx::xcd_tile
    double* out;
    int* status;
};

// Detector epilogue of the fused tile pass (EMCCD + readout noise + ADC of
// the tile's pixels straight from the LDS accumulators).
struct DetectArgs {
    smx_detector_params_t P;
    const double* gain_map;      // NULL: P.gain
    const double* offset_map;    // NULL: P.offset
    uint64_t seed;
    uint32_t frame, pix_base;
    const uint32_t* frame_dev;   // NULL, or added to frame
    double* E_out;               // NULL, or QE * (photons + background)
    void* adc_out;
    int adc_f32;                 // 1: float ADC, 0: double
};


constexpr int FCHUNK = 64;

// General-path pixel integral for one emitter (used for flagged emitters).
// Inlined: a call would need the kernel argument block in private memory.
__device__ __forceinline__ double sat_pixel(const FastArgs& A, const smx_fov_t& F, int e, double w,
                                            int pi, int pj, double acc) {
    int slot = A.key[e];
    Axis au = make_axis(F.half_w, A.u[e] - F.p0u, A.ext_half, F.pixel_length, F.nw);
    Axis av = make_axis(F.half_h, A.v[e] - F.p0v, A.ext_half, F.pixel_length, F.nh);
    if (pi < au.k0 || pi >= au.k1 || pj < av.k0 || pj >= av.k1) return acc;
    int b0 = axis_boundary(au, pi, F.pixel_length, A.res, A.ntab);
    int b1 = axis_boundary(au, pi + 1, F.pixel_length, A.res, A.ntab);
    int c0 = axis_boundary(av, pj, F.pixel_length, A.res, A.ntab);
    int c1 = axis_boundary(av, pj + 1, F.pixel_length, A.res, A.ntab);
    if (b1 <= b0 || c1 <= c0) return acc;
    const double* S = A.sat + (size_t)slot * A.sat_stride;
    double s = ((S[A.L.index(b1, c1)] - S[A.L.index(b0, c1)]) - S[A.L.index(b1, c0)]) +
               S[A.L.index(b0, c0)];
    double photons = s * A.unit_area;
    if (photons > 0.0) acc += photons * w;
    return acc;
}

// One wave per 16 x 16 quadrant of a tile of QPW x QPW quadrants (side 16,
// 32 or 64: the tile shape the batched path's autotuner picks).  Per chunk of
// the sorted tile list, each wave clips the emitters' footprints to its
// quadrant and splits each rectangle into 64-pixel parts (lane = pixel of the
// rectangle, so no lane idles outside it).  The parts go to LDS as decoded
// records, up to PMAX per round, padded with null records to whole groups of
// NSLOT; per group the lanes issue NSLOT table loads at once (24-bit
// multiply-adds on uniform record fields, byte offsets from the table base),
// then add the contributions into the wave's LDS accumulator part by part in
// list order, with no branch: lanes outside a part add into their own spill
// slot.  Every pixel sums its emitters in emitter order, and the lanes of one
// part touch distinct pixels.  KT is the phase-table element type: double
// (values identical to the SAT path) or float (table value times the weight
// in float32: each contribution within ~3 x 2^-24 relative).
// Quadrant side QS (one wave each): 16, or 32.  Accumulator row stride QS + 1
// (doubles).  ds_add_f64 serves 32 lanes per 256 bytes of banks; with 16 the
// rows of a narrow part share banks, 17 spreads them (measured: 17 is the
// fastest of 16, 17, 18, 20 and 24 for 16 x 16 quadrants).
template <int QS> constexpr int acc_stride() { return QS + 1; }
constexpr int NSLOT = 8;          // 64-pixel parts in flight per wave
// tile list sorted in LDS: 2048 entries, 1024 for one-wave tiles
template <int QPW> constexpr int psort_cap() { return QPW == 1 ? 1024 : 2048; }

// One 64-pixel part of an emitter's footprint rectangle inside a quadrant,
// decoded for the lanes: pixel idx = p64 + lane of the rectangle (row-major,
// width W) is valid for lane < lim; row = (idx * magic) >> 16 (exact for
// idx < 1024, W <= 32) and col = idx - row W, so its table element
// kq + row su + col sv is kq + idx sv + row (su - W sv) (bytes) and its
// accumulator slot pix + 8 (row AS + col) is pix + 8 idx + row 8 (AS - W):
// two 24-bit multiply-adds each, col never formed.
struct __align__(16) PartRec {
    int kq;          // table byte offset (32-bit form) or low half of the byte address
    int magic_lim;   // ceil(65536 / W) | lim << 24 (valid lanes, 0..64)
    int dk;          // su - W sv: table bytes per rectangle row beyond idx sv
    int p64;         // first rectangle pixel of the part
    int da;          // 8 (AS - W): accumulator bytes per rectangle row beyond 8 idx
    int sv;          // table bytes per pixel along v
    int pix;         // accumulator byte offset of the rectangle's first pixel
    int kq_hi;       // high half of the byte address (64-bit form)
};
constexpr int PMAX = 64;          // parts per wave per round (one lane writes each padding record)
// waves per SIMD the register allocation leaves room for: 6 (<= 80 VGPRs; a
// 32-pixel tile's block takes 22.5 KB of LDS, so 7 fit a CU; 7 waves would
// spill)
#ifndef SMX_SPLAT_WPE
#define SMX_SPLAT_WPE 6
#endif

// LDS plan of the tile pass (byte offsets into one buffer).  With the
// detector epilogue (DET) the EMCCD work queues of emccd_block take the same
// buffer once the accumulators are read: from offset 0 for tiles of up to
// 32 x 32 pixels (one round), after the accumulators for 64 x 64 tiles
// (four rounds of 1024 pixels that read the accumulators in turn).
template <int QPW, bool DET, int TAPS, int QS>
struct SplatLds {
    static constexpr int AS = acc_stride<QS>();
    static constexpr int WPT = QPW * QPW, TSZ = QS * QPW, ACC = QS * AS;
    // accumulators, then one spill slot per lane shared by the waves (lanes
    // outside a part add there: no branch per part, the slots are never read);
    // before the accumulators are cleared the same space sorts the tile list
    static constexpr int ACCN = WPT * ACC + 64 > psort_cap<QPW>() / 2 ? WPT * ACC + 64
                                                                       : psort_cap<QPW>() / 2;
    // the merge sort's runs and output (up to 4 entries per thread each, at
    // most the sort capacity) fit too
    static_assert(WPT == 1 || 2 * (4 * 64 * WPT < psort_cap<QPW>() ? 4 * 64 * WPT : psort_cap<QPW>()) <= 2 * ACCN,
                  "merge sort LDS");
    static constexpr int GEO = (ACCN * 8 + 15) & ~15;   // int4 lo_u, lo_v (frame pixels), n_u, n_v
    static constexpr int STR = GEO + FCHUNK * 16;       // int2 table strides per pixel along u, v
    static constexpr int FLAGS = STR + FCHUNK * 8;
    static constexpr int EID = FLAGS + FCHUNK * 4;
    static constexpr int W = EID + FCHUNK * 4;          // double emitter weights
    static constexpr int KO = W + FCHUNK * 8;           // table element of footprint pixel (0, 0)
    static constexpr int PR = KO + FCHUNK * 8;          // PartRec [WPT][PMAX]: parts of one round
    static constexpr int PW = PR + WPT * PMAX * 32;     // [WPT][PMAX] weights: double, or
                                                        // float4 tap coefficients (TAPS 4)
    static constexpr int FRAC = PW + WPT * PMAX * (TAPS == 4 ? 16 : 8);   // LUT fractions
    static constexpr int SPLAT_END = FRAC + (TAPS == 4 ? FCHUNK * 4 : 0);
    // detector epilogue: rounds of RP pixels, PPT per thread (at most 4)
    static constexpr int NT = 64 * WPT;
    static constexpr int RPMAX = 4 * NT < 1024 ? 4 * NT : 1024;
    static constexpr int RP = TSZ * TSZ < RPMAX ? TSZ * TSZ : RPMAX;
    static constexpr int PPT = RP / NT;
    static constexpr int ROUNDS = TSZ * TSZ / RP;
    static constexpr int QOFF = ROUNDS == 1 ? 0 : GEO;
    static constexpr int Q_END = QOFF + smx::emccd::lds_bytes(RP);
    static constexpr int BYTES = DET && Q_END > SPLAT_END ? Q_END : SPLAT_END;
};

template <typename KT, int QPW, bool O32, bool DET, int TAPS, int QS>
__global__ void __launch_bounds__(64 * QPW * QPW) __attribute__((amdgpu_waves_per_eu(SMX_SPLAT_WPE, 8)))
splat_phase_kernel(FastArgs A, DetectArgs D) {
    using LP = SplatLds<QPW, DET, TAPS, QS>;
    constexpr int AS = LP::AS;
    // 64-pixel parts in flight per group: NSLOT table loads (one 16-byte cell
    // entry per part in LUT mode)
    constexpr int NS = NSLOT;
    constexpr int WPT = LP::WPT, TSZ = LP::TSZ, PSORT_CAP = psort_cap<QPW>();
    constexpr int ACC = LP::ACC, ACCN = LP::ACCN;
    __shared__ __align__(16) char s_lds[LP::BYTES];
    double* const s_acc = (double*)s_lds;
    int* const s_sort = (int*)s_acc;
    int4* const s_geo = (int4*)(s_lds + LP::GEO);
    int2* const s_str = (int2*)(s_lds + LP::STR);
    int* const s_flags = (int*)(s_lds + LP::FLAGS);
    int* const s_e = (int*)(s_lds + LP::EID);
    double* const s_w = (double*)(s_lds + LP::W);
    long long* const s_ko = (long long*)(s_lds + LP::KO);
    PartRec (*const s_pr)[PMAX] = (PartRec (*)[PMAX])(s_lds + LP::PR);
    double (*const s_pw)[PMAX] = (double (*)[PMAX])(s_lds + LP::PW);
    float4 (*const s_pc)[PMAX] = (float4 (*)[PMAX])(s_lds + LP::PW);
    unsigned* const s_frac = (unsigned*)(s_lds + LP::FRAC);
    (void)ACCN;

    const int tile = A.xcd ? smx::xcd_tile(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    const int fov = A.tile_fov[tile];
    if (fov < 0 || fov >= A.nfov) return;
    const smx_fov_t F = A.fovs[fov];
    const int local = tile - F.tile_off;
    const int tr = local / F.tiles_c, tc = local % F.tiles_c;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int qr0 = tr * TSZ + (wave / QPW) * QS, qc0 = tc * TSZ + (wave % QPW) * QS;
    const int qh = min(QS, F.nw - qr0), qw = min(QS, F.nh - qc0);
    const bool active = qh > 0 && qw > 0;
    double* acc = s_acc + wave * ACC;
    const char* kbase = (const char*)A.Kraw;
    const unsigned lane24 = ((unsigned)lane << 24) | 0xffffffu;   // lane < lim <=> lane24 < magic_lim
    const int acc_off = wave * (ACC * 8);                          // this wave's accumulator (bytes)
    const unsigned spill = (unsigned)((WPT * ACC + lane) * 8);

    const int begin = A.tile_begin[tile];
    int count = A.tile_begin[tile + 1] - begin;
    if ((long long)begin + count > A.capacity) count = (int)max(0LL, A.capacity - begin);
    int* list = A.list + begin;
    const bool sorted = count <= PSORT_CAP;
    // blocks of several waves: up to one entry per thread, rank sort by the
    // whole block (its input staged after the output); up to 4 per thread,
    // every wave sorts a run in registers and the runs are ranked against each
    // other (merge_rank_sort: measured, the rank sort is the faster of the two
    // for the ~80-entry lists of 32 px tiles, the merge for longer ones)
    constexpr int RANK_MAX = WPT > 1 ? 64 * WPT : 0;
    if (count >= 2 && count <= RANK_MAX) {
        smx::rank_sort(list, count, s_sort + RANK_MAX, s_sort);
    } else if (count >= 2 && count <= 2 * RANK_MAX) {
        smx::merge_rank_sort<2>(list, count, s_sort + 2 * RANK_MAX, s_sort);
    } else if (count >= 2 && count <= 4 * RANK_MAX && count <= PSORT_CAP) {
        smx::merge_rank_sort<4>(list, count, s_sort + 4 * RANK_MAX, s_sort);
    } else if (count >= 2 && count <= WAVE_SORT_MAX) {
        // list of up to 1024 entries: one wave sorts it in registers
        if (wave == 0) wave_sort_list(list, count, lane, s_sort);
        __syncthreads();
    } else if (count >= 2 && sorted) {
        int np2 = 1;
        while (np2 < count) np2 <<= 1;
        for (int k = threadIdx.x; k < np2; k += blockDim.x) s_sort[k] = k < count ? list[k] : 0x7fffffff;
        __syncthreads();
        lds_bitonic_sort(s_sort, np2);
    }
    if (count >= 2 && sorted) {
        // sorted list back in place (read per chunk below), then the space
        // becomes the accumulators
        for (int k = threadIdx.x; k < count; k += blockDim.x) list[k] = s_sort[k];
        __syncthreads();
    }
    for (int k = lane; k < QS * AS; k += 64) acc[k] = 0.0;

    for (int base = 0; base < count; base += FCHUNK) {
        const int nc = min(FCHUNK, count - base);
        if (threadIdx.x < nc) {
            int e = smx::gload(list + base + threadIdx.x);
            EmitterRec r = A.rec[e];
            s_e[threadIdx.x] = e;
            s_geo[threadIdx.x] = make_int4(r.lo_u, r.lo_v, r.n_u, r.n_v);
            s_str[threadIdx.x] = make_int2(r.su, r.sv);
            s_flags[threadIdx.x] = r.flags;
            s_w[threadIdx.x] = r.w;
            if (TAPS == 4) s_frac[threadIdx.x] = r.frac;
            s_ko[threadIdx.x] = r.flags == 0 ? r.table_off : 0;
            if (r.flags & 2) atomicOr(A.status, 8);
        }
        __syncthreads();
        if (active) {
            // footprint rectangle of chunk emitter `lane` inside this quadrant:
            // origin (rr0, cc0) in the quadrant, width W, `area` pixels in np
            // parts, kq = table element of its first pixel
            int np = 0, W = 1, area = 0, rr0 = 0, cc0 = 0, su = 0, sv = 0;
            long long kq = 0;
            bool irr = false;
            if (lane < nc && s_flags[lane] != 2) {
                const int4 g = s_geo[lane];
                const int r0 = max(g.x, qr0), r1 = min(g.x + g.z, qr0 + qh);
                const int c0 = max(g.y, qc0), c1 = min(g.y + g.w, qc0 + qw);
                if (r1 > r0 && c1 > c0) {
                    const int2 st = s_str[lane];
                    W = c1 - c0;
                    area = (r1 - r0) * W;
                    np = (area + 63) >> 6;
                    irr = s_flags[lane] == 1;
                    rr0 = r0 - qr0;
                    cc0 = c0 - qc0;
                    su = st.x;
                    sv = st.y;
                    kq = s_ko[lane] + (long long)(r0 - g.x) * su + (long long)(c0 - g.y) * sv;
                }
            }
            const bool any_irregular = __ballot(irr) != 0;
            const int incl = smx::wave_incl_scan(np);
            const int first = incl - np;   // this emitter's parts: [first, incl) in list order
            const int total = __builtin_amdgcn_readlane(incl, 63);

            if (any_irregular) {
                // rare (a float64 tie in a cell boundary): one part at a time, in
                // order, flagged emitters integrated through the SAT
                for (int b = 0; b < total; ++b) {
                    const int c = __popcll(__ballot(incl <= b));   // emitter lane owning part b
                    const int p = b - __shfl(first, c);
                    const int Wc = __shfl(W, c), ar = __shfl(area, c);
                    const int rc = __shfl(rr0, c), cc = __shfl(cc0, c);
                    const int suc = __shfl(su, c), svc = __shfl(sv, c);
                    const long long kqc = __shfl(kq, c);
                    const int idx = p * 64 + lane;
                    if (idx >= ar) continue;
                    const int row = idx / Wc, col = idx - row * Wc;
                    double val;
                    if (s_flags[c] == 1) {
                        val = sat_pixel(A, F, s_e[c], s_w[c], qr0 + rc + row, qc0 + cc + col, 0.0);
                    } else {
                        val = (double)smx::gload((const KT*)A.Kraw + kqc + (long long)row * suc +
                                                 (long long)col * svc) *
                              s_w[c];
                    }
                    lds_add(acc + (rc + row) * AS + cc + col, val);
                }
            } else {
                PartRec* pr = s_pr[wave];
                double* pw = s_pw[wave];
                // float tables: the weight in float32 too, the product converted
                // to float64 for the add (contribution within ~3 2^-24 relative)
                constexpr bool F32W = std::is_same<KT, float>::value && TAPS == 1;
                float* pwf = reinterpret_cast<float*>(pw);
                float4* pc = s_pc[wave];
                const int magic = (65536 + W - 1) / W;
                const double w = lane < nc ? s_w[lane] : 0.0;
                // LUT mode: bilinear weights of the four phase taps, times w
                float4 cw = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (TAPS == 4 && lane < nc) {
                    const unsigned fr = s_frac[lane];
                    const float fu = (float)(fr & 0xffffu) * (1.0f / 65535.0f);
                    const float fv = (float)(fr >> 16) * (1.0f / 65535.0f);
                    const float wf = (float)w;
                    cw = make_float4(((1.0f - fu) * (1.0f - fv)) * wf, ((1.0f - fu) * fv) * wf,
                                     (fu * (1.0f - fv)) * wf, (fu * fv) * wf);
                }
                const long long kb = kq * (long long)sizeof(KT);
                const unsigned long long ka = (unsigned long long)(kbase + kb);
                const int pix = acc_off + (rr0 * AS + cc0) * 8;
                for (int rb = 0; rb < total; rb += PMAX) {
                    // this round's parts [rb, rb + PMAX), then null records up to
                    // a whole number of NSLOT groups (no lane valid, table offset 0)
                    for (int p = 0; p < np; ++p) {
                        const int pos = first + p - rb;
                        if ((unsigned)pos >= (unsigned)PMAX) continue;
                        const int lim = min(64, area - p * 64);
                        PartRec& R = pr[pos];
                        R.kq = O32 ? (int)(unsigned)kb : (int)(unsigned)ka;
                        R.magic_lim = magic | (lim << 24);
                        R.dk = (su - W * sv) * (int)sizeof(KT);
                        R.p64 = p * 64;
                        R.da = (AS - W) * 8;
                        R.sv = sv * (int)sizeof(KT);
                        R.pix = pix;
                        R.kq_hi = O32 ? 0 : (int)(unsigned)(ka >> 32);
                        if (TAPS == 4) pc[pos] = cw;
                        else if (F32W) pwf[pos] = (float)w;
                        else pw[pos] = w;
                    }
                    const int cnt = min(PMAX, total - rb);
                    const int cntp = (cnt + NS - 1) & ~(NS - 1);
                    if (lane >= cnt && lane < cntp) {
                        PartRec& R = pr[lane];
                        R.kq = O32 ? 0 : (int)(unsigned)(unsigned long long)kbase;
                        R.magic_lim = 0;
                        R.dk = 0;
                        R.p64 = 0;
                        R.da = 0;
                        R.sv = 0;
                        R.pix = 0;
                        R.kq_hi = O32 ? 0 : (int)(unsigned)((unsigned long long)kbase >> 32);
                        if (TAPS == 4) pc[lane] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                        else if (F32W) pwf[lane] = 0.0f;
                        else pw[lane] = 0.0;
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    for (int b = 0; b < cntp; b += NS) {
                        KT kv[NS];
                        float4 kc[NS];   // LUT cell: the four taps of a pixel
                        double wv[NS];
                        float wf[NS];
                        unsigned la[NS];
#pragma unroll
                        for (int s = 0; s < NS; ++s) {
                            const PartRec& R = pr[b + s];
                            const int4 h = *(const int4*)&R.kq;   // kq, magic_lim, dk, p64
                            const int4 g = *(const int4*)&R.da;   // da, sv, pix, kq_hi
                            if (F32W) wf[s] = pwf[b + s];
                            else if (TAPS == 1) wv[s] = pw[b + s];
                            const unsigned idx = (unsigned)lane + (unsigned)h.w;
                            const bool ok = lane24 < (unsigned)h.y;
                            const int row = (int)(__umul24(idx, (unsigned)h.y) >> 16);
                            const char* kp;
                            if (O32) {
                                int off = __mul24(row, h.z) + (__mul24((int)idx, g.y) + h.x);
                                off = ok ? off : 0;
                                kp = kbase + (unsigned)off;
                            } else {
                                int off = __mul24(row, h.z) + __mul24((int)idx, g.y);
                                off = ok ? off : 0;
                                kp = (const char*)(((unsigned long long)(unsigned)g.w << 32) |
                                                   (unsigned)h.x) + off;
                            }
                            if (TAPS == 4) kc[s] = smx::gload4((const float*)kp);
                            else kv[s] = smx::gload((const KT*)kp);
                            const unsigned l = (unsigned)__mul24(row, g.x) + ((unsigned)g.z + (idx << 3));
                            la[s] = ok ? l : spill;
                        }
#pragma unroll
                        for (int s = 0; s < NS; ++s) {
                            if (TAPS == 4) {
                                const float4 c = pc[b + s];
                                const float val = ((c.x * kc[s].x + c.y * kc[s].y) + c.z * kc[s].z) +
                                                  c.w * kc[s].w;
                                lds_add((double*)((char*)s_acc + la[s]), (double)val);
                            } else {
                                lds_add((double*)((char*)s_acc + la[s]),
                                        F32W ? (double)((float)kv[s] * wf[s]) : (double)kv[s] * wv[s]);
                            }
                        }
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                }
            }
        }
        __syncthreads();
    }
    if (!DET) {
        if (active) {
            for (int k = lane; k < QS * QS; k += 64) {
                const int r = k / QS, c = k % QS;
                const double a = acc[r * AS + c];
                if (r >= qh || c >= qw) continue;
                const size_t o = (size_t)F.pix_off + (size_t)(qr0 + r) * F.nh + (qc0 + c);
                if (A.accumulate) {
                    if (count > 0) A.out[o] = A.out[o] + a;
                } else {
                    // streaming store: keep the phase table, not the frame, in L2
                    __builtin_nontemporal_store(a, A.out + o);
                }
            }
        }
        return;
    }
    // Detector epilogue: the tile's photons go from the accumulators through
    // the EMCCD register, readout noise and ADC (_epifm.py:1436-1469) without
    // a round trip through HBM; the draws are the standalone detector's
    // (emccd_block, same Philox stream per pixel).
    __syncthreads();   // every quadrant's accumulators complete (also when count == 0)
    constexpr int NT = LP::NT, RP = LP::RP, PPT = LP::PPT;
    const smx::emccd::BlockLds L = smx::emccd::carve(s_lds + LP::QOFF, RP);
    const uint32_t frame = D.frame + (D.frame_dev ? *D.frame_dev : 0u);
    const uint32_t k0 = (uint32_t)D.seed, k1 = (uint32_t)(D.seed >> 32);
    const int t = threadIdx.x;
    const int row0 = tr * TSZ, col0 = tc * TSZ;
    // the tile's first pixel in the batch buffer; a pixel's offset from it,
    // r * nh + c, stays 32-bit (a tile spans at most 64 frame rows)
    const long long tbase = F.pix_off + (long long)row0 * F.nh + col0;
    const uint32_t pid0 = D.pix_base + (uint32_t)tbase;
    for (int rd = 0; rd < LP::ROUNDS; ++rd) {
        float E[PPT];
        bool in[PPT];
        int o[PPT];
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
            const int p = rd * RP + t + j * NT;
            const int r = p / TSZ, c = p % TSZ;
            in[j] = row0 + r < F.nw && col0 + c < F.nh;
            o[j] = r * F.nh + c;
            E[j] = 0.0f;
            const double a = s_acc[((r / QS) * QPW + (c / QS)) * ACC + (r % QS) * AS + (c % QS)];
            if (in[j]) {
                E[j] = smx::emccd::expectation_f32(a, D.P);
                if (A.out) __builtin_nontemporal_store(a, A.out + tbase + o[j]);
                if (D.E_out) D.E_out[tbase + o[j]] = (double)E[j];
            }
        }
        if (LP::QOFF == 0) __syncthreads();   // accumulators read before the queues overlay them
        float S[PPT], noise[PPT];
        smx::emccd::emccd_block<NT, PPT>(
                L, E, in,
                [&](int loc) {
                    const int p = rd * RP + loc;
                    return pid0 + (uint32_t)((p / TSZ) * F.nh + (p % TSZ));
                },
                frame, k0, k1, D.P, S, noise);
        float* const adc32 = (float*)D.adc_out + tbase;
        double* const adc64 = (double*)D.adc_out + tbase;
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
            if (!in[j]) continue;
            const float gain = (float)(D.gain_map ? D.gain_map[tbase + o[j]] : D.P.gain);
            const float offset = (float)(D.offset_map ? D.offset_map[tbase + o[j]] : D.P.offset);
            const float x = smx::emccd::adc_f32(S[j], noise[j], D.P, gain, offset);
            if (D.adc_f32) adc32[o[j]] = x;
            else adc64[o[j]] = (double)x;
        }
    }
}

struct SplatWorkspace {
    EmitterRec* rec;   // n (phase-table path)
    int* tile_count;   // (n_tiles + 1) x COUNT_STRIDE (padded counters)
    int* tile_dense;   // n_tiles + 1
    int* tile_begin;   // n_tiles + 1
    int* tile_fov;     // n_tiles
    int* list;         // capacity
    int* long_tiles;   // n_tiles + 1: tiles with lists too long to sort in LDS
    void* cub_tmp;
    size_t cub_bytes;
};

size_t scan_tmp_bytes(int n_tiles) {
    size_t bytes = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (int*)nullptr, (int*)nullptr, n_tiles + 1);
    return bytes;
}

size_t layout_workspace(int n, int n_tiles, long long capacity, char* base, SplatWorkspace* ws) {
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += align256(bytes); return o; };
    size_t o_rec = take(sizeof(EmitterRec) * (size_t)(n > 0 ? n : 1));
    size_t o_count = take(sizeof(int) * (size_t)(n_tiles + 1) * smx::COUNT_STRIDE);
    size_t o_dense = take(sizeof(int) * (n_tiles + 1));
    size_t o_begin = take(sizeof(int) * (n_tiles + 1));
    size_t o_fov = take(sizeof(int) * n_tiles);
    size_t o_list = take(sizeof(int) * (size_t)(capacity > 0 ? capacity : 1));
    size_t o_long = take(sizeof(int) * (size_t)(n_tiles + 1));
    size_t cb = scan_tmp_bytes(n_tiles);
    size_t o_cub = take(cb);
    if (ws && base) {
        ws->rec = (EmitterRec*)(base + o_rec);
        ws->tile_count = (int*)(base + o_count);
        ws->tile_dense = (int*)(base + o_dense);
        ws->tile_begin = (int*)(base + o_begin);
        ws->tile_fov = (int*)(base + o_fov);
        ws->list = (int*)(base + o_list);
        ws->long_tiles = (int*)(base + o_long);
        ws->cub_tmp = base + o_cub;
        ws->cub_bytes = cb;
    }
    return off;
}

}  // namespace

extern "C" {

int smx_emitter_weights(int n, const double* depth, const double* p_state,
                        const smx_weight_params_t* params, double unit_time, double* budget,
                        const int* budget_slot, const int* slot_map, int n_slot_map,
                        double* w_out, int* key_out, double* pstate_out, hipStream_t stream) {
    if (n < 0 || !params || (n > 0 && (!depth || !w_out || !key_out))) {
        smx::set_error("smx_emitter_weights: invalid arguments (n=%d)", n);
        return SMX_EVALUE;
    }
    if (n == 0) return SMX_OK;
    hipLaunchKernelGGL(weights_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, n, depth,
                       p_state, *params, unit_time, budget, budget_slot, slot_map, n_slot_map,
                       w_out, key_out, pstate_out);
    SMX_CHECK_LAUNCH("weights_kernel");
    return SMX_OK;
}

int smx_splat_tiles(const smx_fov_t* fov, int tile) {
    (void)tile;
    if (!fov) return 0;
    return ((fov->nw + TS - 1) / TS) * ((fov->nh + TS - 1) / TS);
}

int smx_splat_tile_size(void) { return TS; }

size_t smx_splat_workspace_bytes(int n_emitters, int n_tiles, long long list_capacity) {
    if (n_emitters < 0 || n_tiles <= 0 || list_capacity < 0) return 0;
    return layout_workspace(n_emitters, n_tiles, list_capacity, nullptr, nullptr);
}

size_t smx_phase_table_elems(int period, int m) {
    if (period < 1 || m < 1) return 0;
    return (size_t)period * period * m * m;
}

int smx_phase_table_to_float(const double* K, size_t n, float* K32, hipStream_t stream);

int smx_phase_table_build(const double* sat, size_t sat_stride, int nslots, int ntab,
                          int sat_period, int m, double unit_area, double* K, hipStream_t stream) {
    if (!sat || !K || nslots < 0 || ntab < 1 || sat_period < 2 || m < 1) {
        smx::set_error("smx_phase_table_build: invalid arguments (period=%d, m=%d)", sat_period, m);
        return SMX_EVALUE;
    }
    if (nslots == 0) return SMX_OK;
    size_t total = smx_phase_table_elems(sat_period, m) * (size_t)nslots;
    smx::SatLayout L = smx::make_sat_layout(ntab + 1, sat_period);
    hipLaunchKernelGGL(phase_table_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       stream, sat, sat_stride, L, ntab, sat_period, m, nslots, unit_area, K);
    SMX_CHECK_LAUNCH("phase_table_kernel");
    return SMX_OK;
}

// Tiles of one frame share emitters and table blocks.  With many frames,
// whole frames per XCD (smx::xcd_tile) keep that sharing in one L2 (headline
// batch: tile pass -2%); one large frame does better with the dispatcher's
// round robin, whose neighbouring tiles run at the same time on all XCDs and
// meet their shared lines in the memory-side cache (1024^2 LUT frame: 0.63 ms
// round robin, 0.76 ms banded).
static int xcd_map(int nfov) { return nfov >= smx::NXCD ? 1 : 0; }

// stages: 1 = bin the emitters to tiles (and, with K, their table records),
// 2 = accumulate the tiles from the binned workspace, 3 = both.
static int splat_run(int n, const double* u, const double* v, const double* w, const int* key,
                     const int* fov_id, const smx_fov_t* fovs_dev, int nfov, int n_tiles,
                     const double* sat, size_t sat_stride, int n_slots, int ntab, int sat_period,
                     double res, double unit_area, long long list_capacity, int accumulate,
                     const void* K, int k_format, int m, double* out, void* workspace,
                     size_t workspace_bytes, int* status_dev, hipStream_t stream, int stages,
                     int tile_side, const DetectArgs* det = nullptr,
                     const smx_lut_t* lut = nullptr) {
    LutRec LR{0, 0, 0, 0};
    if (lut) {
        // LUT mode: float32 pixel-integral tables per depth slice, four taps
        if (!K || lut->q < 1 || lut->m != 2 * lut->c + 1 || lut->c < 1 || lut->nslices < 1) {
            smx::set_error("smx_splat_lut: invalid table description");
            return SMX_EVALUE;
        }
        LR.q = lut->q;
        LR.m = lut->m;
        LR.c = lut->c;
        LR.slice_elems = (size_t)lut->q * lut->q * lut->m * lut->m * 4;   // cell form
        sat = (const double*)K;    // not read: LUT records are never integrated through a SAT
        sat_stride = 0;
        n_slots = lut->nslices;
        ntab = 1;
        sat_period = 2;
        m = lut->m;
        k_format = 1;
    }
    const smx::PhaseFold PF = smx::make_phase_fold(ntab, sat_period > 0 ? sat_period : 1);
    const bool q32 = (tile_side & SMX_TILE_Q32) != 0;
    const bool warm = (tile_side & SMX_TILE_WARM) != 0;
    const int side = tile_side & ~(SMX_TILE_Q32 | SMX_TILE_WARM);
    const int lts = side == 16 ? 4 : side == 32 ? 5 : side == 64 ? 6 : -1;
    if (lts < 0 || (!K && (lts != 5 || q32)) || (q32 && lts == 4)) {
        smx::set_error("smx_splat_expected: tile side %d (16, 32 or 64 with phase tables, "
                       "SMX_TILE_Q32 with 32 or 64; 32 without)", tile_side);
        return SMX_EVALUE;
    }
    if (n < 0 || nfov <= 0 || n_tiles <= 0 || ntab < 1 || sat_period < 1 || !fovs_dev ||
        (!out && !det) ||
        (K && (sat_period < 2 || m < 1 || k_format < 0 || k_format > 2)) ||
        (K && k_format == 2 && PF.mf > m) ||
        !workspace || !status_dev || (n > 0 && (!u || !v || !w || !key || !sat))) {
        smx::set_error("smx_splat_expected: invalid arguments (n=%d, nfov=%d, n_tiles=%d)", n,
                       nfov, n_tiles);
        return SMX_EVALUE;
    }
    size_t need = layout_workspace(n, n_tiles, list_capacity, nullptr, nullptr);
    if (workspace_bytes < need) {
        smx::set_error("smx_splat_expected: workspace too small (%zu < %zu bytes)",
                       workspace_bytes, need);
        return SMX_EVALUE;
    }
    SplatWorkspace ws;
    layout_workspace(n, n_tiles, list_capacity, (char*)workspace, &ws);
    const double ext_half = (res * (double)(ntab - 1)) * 0.5;   // signal_width * 0.5
    const int nb = (n + 255) / 256;
    const int nbr = (n + REC_THREADS - 1) / REC_THREADS;
    // lists the tile pass sorts itself; longer ones are sorted after binning
    const int long_min = !K ? SORT_CAP : (lts == 4 || (lts == 5 && q32)) ? psort_cap<1>()
                                                                         : psort_cap<2>();
    // about two tiles per emitter at these footprints; the record path bins
    // with per-block aggregation, the SAT path with one atomic per entry
    const int cs = K ? smx::agg_count_stride(2LL * n, n_tiles) : smx::count_stride(2LL * n, n_tiles);

    if (stages & 1) {
        // a warm workspace's counters are at zero and its tile map current;
        // the one-workgroup scan clears the long-list count itself
        if (!(warm && n_tiles + 1 <= smx::SCAN_MAX)) {
            hipLaunchKernelGGL(smx::tiles_init_kernel, dim3((n_tiles + 1 + 255) / 256), dim3(256),
                               0, stream, nfov, fovs_dev, n_tiles, ws.tile_count, cs, ws.tile_fov,
                               ws.long_tiles);
            SMX_CHECK_LAUNCH("tiles_init_kernel");
        }
        if (n > 0) {
            if (K) {
                // records first; they count their tiles as they are written
                const int folded = k_format == 2;
                hipLaunchKernelGGL(emitter_rec_kernel, dim3(nbr), dim3(REC_THREADS), 0, stream, n,
                                   u, v, w, key, fov_id, fovs_dev, nfov, ext_half, res, ntab,
                                   sat_period, m, n_slots,
                                   folded ? PF.elems() : smx_phase_table_elems(sat_period, m),
                                   folded, PF, ws.rec, n_tiles, ws.tile_count, cs, status_dev,
                                   lts, LR);
                SMX_CHECK_LAUNCH("emitter_rec_kernel");
            } else {
                hipLaunchKernelGGL(bin_count_kernel, dim3(nb), dim3(256), 0, stream, n, u, v, w,
                                   fov_id, fovs_dev, nfov, ext_half, n_tiles, ws.tile_count, cs);
                SMX_CHECK_LAUNCH("bin_count_kernel");
            }
        }
        if (n_tiles + 1 <= smx::SCAN_MAX) {
            // one workgroup scans the counters in place of compact + device scan
            hipLaunchKernelGGL(smx::scan_counts_kernel, dim3(1), dim3(smx::SCAN_THREADS), 0, stream,
                               ws.tile_count, cs, n_tiles + 1, ws.tile_begin, long_min,
                               ws.long_tiles);
            SMX_CHECK_LAUNCH("scan_counts_kernel");
        } else {
            size_t cb = ws.cub_bytes;
            hipLaunchKernelGGL(smx::compact_counts_kernel, dim3((n_tiles + 1 + 255) / 256), dim3(256),
                               0, stream, ws.tile_count, cs, n_tiles + 1, ws.tile_dense, long_min,
                               ws.long_tiles);
            SMX_CHECK_LAUNCH("compact_counts_kernel");
            if (smx::check_hip(hipcub::DeviceScan::ExclusiveSum(ws.cub_tmp, cb, ws.tile_dense,
                                                                ws.tile_begin, n_tiles + 1, stream),
                               "hipcub ExclusiveSum"))
                return SMX_ERUNTIME;
        }
        if (n > 0) {
            if (K)
                hipLaunchKernelGGL(bin_fill_rec_kernel, dim3(nbr), dim3(REC_THREADS), 0, stream, n,
                                   ws.rec, fov_id, fovs_dev, nfov, n_tiles, ws.tile_begin,
                                   ws.tile_count, cs, ws.list, list_capacity, status_dev, lts);
            else
                hipLaunchKernelGGL(bin_fill_kernel, dim3(nb), dim3(256), 0, stream, n, u, v, w,
                                   fov_id, fovs_dev, nfov, ext_half, n_tiles, ws.tile_begin,
                                   ws.tile_count, cs, ws.list, list_capacity, status_dev);
            SMX_CHECK_LAUNCH("bin_fill_kernel");
            hipLaunchKernelGGL(smx::sort_long_lists_kernel, dim3(smx::LONG_SORT_BLOCKS),
                               dim3(smx::LONG_SORT_THREADS), 0, stream, ws.long_tiles,
                               ws.tile_begin, ws.list, list_capacity, status_dev);
            SMX_CHECK_LAUNCH("sort_long_lists_kernel");
        }
    }
    if (!(stages & 2)) return SMX_OK;
    if (K) {
        FastArgs B;
        B.Kraw = K;
        B.u = u; B.v = v; B.key = key; B.fovs = fovs_dev; B.nfov = nfov;
        B.tile_fov = ws.tile_fov; B.tile_begin = ws.tile_begin; B.list = ws.list;
        B.capacity = list_capacity; B.rec = ws.rec; B.K = (const double*)K; B.M = m;
        B.sat = sat; B.sat_stride = sat_stride; B.L = smx::make_sat_layout(ntab + 1, sat_period);
        B.res = res; B.ext_half = ext_half; B.unit_area = unit_area; B.ntab = ntab;
        B.accumulate = accumulate; B.out = out; B.status = status_dev; B.xcd = xcd_map(nfov);
        const dim3 grid((unsigned)n_tiles);
        // table byte offsets fit 32 bits: parts address it from its base
        const size_t slot_elems = LR.q ? LR.slice_elems
                                : k_format == 2 ? PF.elems() : smx_phase_table_elems(sat_period, m);
        const double table_bytes = (double)n_slots * (double)slot_elems * (k_format != 0 ? 4.0 : 8.0);
        // SMX_SPLAT_ADDR64=1 takes the 64-bit form for any table (tests of that form)
        const char* a64 = getenv("SMX_SPLAT_ADDR64");
        const bool o32 = table_bytes < 4294967296.0 && !(a64 && a64[0] == '1');
        DetectArgs Dv{};
        if (det) Dv = *det;
        auto launch = [&](auto kt, auto o, auto dt, auto taps) {
            using KT = decltype(kt);
            constexpr bool O = decltype(o)::value, DT = decltype(dt)::value;
            constexpr int TP = decltype(taps)::value;
            const bool wide = q32;
            if (lts == 4)
                hipLaunchKernelGGL((splat_phase_kernel<KT, 1, O, DT, TP, 16>), grid, dim3(64), 0, stream, B, Dv);
            else if (lts == 5 && wide)
                hipLaunchKernelGGL((splat_phase_kernel<KT, 1, O, DT, TP, 32>), grid, dim3(64), 0, stream, B, Dv);
            else if (lts == 5)
                hipLaunchKernelGGL((splat_phase_kernel<KT, 2, O, DT, TP, 16>), grid, dim3(256), 0, stream, B, Dv);
            else if (wide)
                hipLaunchKernelGGL((splat_phase_kernel<KT, 2, O, DT, TP, 32>), grid, dim3(256), 0, stream, B, Dv);
            else
                hipLaunchKernelGGL((splat_phase_kernel<KT, 4, O, DT, TP, 16>), grid, dim3(1024), 0, stream, B, Dv);
        };
        using One = std::integral_constant<int, 1>;
        using Four = std::integral_constant<int, 4>;
        auto launch_d = [&](auto kt, auto o, auto taps) {
            if (det) launch(kt, o, std::true_type(), taps);
            else launch(kt, o, std::false_type(), taps);
        };
        if (LR.q) {
            if (o32) launch_d(0.0f, std::true_type(), Four());
            else launch_d(0.0f, std::false_type(), Four());
        } else if (k_format != 0) {
            if (o32) launch_d(0.0f, std::true_type(), One());
            else launch_d(0.0f, std::false_type(), One());
        } else {
            if (o32) launch_d(0.0, std::true_type(), One());
            else launch_d(0.0, std::false_type(), One());
        }
        SMX_CHECK_LAUNCH("splat_phase_kernel");
        return SMX_OK;
    }
    TileArgs A;
    A.u = u; A.v = v; A.w = w; A.key = key; A.fov_id = fov_id; A.fovs = fovs_dev; A.nfov = nfov;
    A.tile_fov = ws.tile_fov; A.tile_begin = ws.tile_begin; A.list = ws.list;
    A.capacity = list_capacity;
    A.sat = sat; A.sat_stride = sat_stride; A.n_slots = n_slots; A.L = smx::make_sat_layout(ntab + 1, sat_period);
    A.res = res; A.ext_half = ext_half; A.unit_area = unit_area; A.ntab = ntab;
    A.accumulate = accumulate; A.out = out; A.status = status_dev; A.xcd = xcd_map(nfov);
    hipLaunchKernelGGL(splat_tile_kernel, dim3(n_tiles), dim3(256), 0, stream, A);
    SMX_CHECK_LAUNCH("splat_tile_kernel");
    return SMX_OK;
}

#define SMX_SPLAT_PARAMS                                                                          \
    int n, const double *u, const double *v, const double *w, const int *key, const int *fov_id, \
        const smx_fov_t *fovs_dev, int nfov, int n_tiles, const double *sat, size_t sat_stride,   \
        int n_slots, int ntab, int sat_period, double res, double unit_area,                      \
        long long list_capacity, int accumulate, const void *K, int k_format, int m, double *out,  \
        void *workspace, size_t workspace_bytes, int *status_dev, hipStream_t stream
#define SMX_SPLAT_ARGS                                                                       \
    n, u, v, w, key, fov_id, fovs_dev, nfov, n_tiles, sat, sat_stride, n_slots, ntab,         \
        sat_period, res, unit_area, list_capacity, accumulate, K, k_format, m, out, workspace, \
        workspace_bytes, status_dev, stream

int smx_splat_expected(SMX_SPLAT_PARAMS) { return splat_run(SMX_SPLAT_ARGS, 3, 32); }
int smx_splat_bin(SMX_SPLAT_PARAMS) { return splat_run(SMX_SPLAT_ARGS, 1, 32); }
int smx_splat_accumulate(SMX_SPLAT_PARAMS) { return splat_run(SMX_SPLAT_ARGS, 2, 32); }
int smx_splat_expected_ts(SMX_SPLAT_PARAMS, int tile_side) {
    return splat_run(SMX_SPLAT_ARGS, 3, tile_side);
}
int smx_splat_bin_ts(SMX_SPLAT_PARAMS, int tile_side) {
    return splat_run(SMX_SPLAT_ARGS, 1, tile_side);
}
int smx_splat_accumulate_ts(SMX_SPLAT_PARAMS, int tile_side) {
    return splat_run(SMX_SPLAT_ARGS, 2, tile_side);
}

static int splat_detect_run(SMX_SPLAT_PARAMS, const double* gain_map, const double* offset_map,
                            const smx_detector_params_t* params, uint64_t seed, uint32_t frame,
                            uint32_t pix_base, const uint32_t* frame_dev, double* E_out,
                            void* adc_out, int adc_is_float32, int tile_side, int stages) {
    if (!K || accumulate != 0 || !params || params->emgain <= 0 || !adc_out) {
        smx::set_error("smx_splat_detect: needs phase tables (K), accumulate = 0, detector "
                       "parameters with emgain > 0 and an ADC buffer");
        return SMX_EVALUE;
    }
    DetectArgs d;
    d.P = *params;
    d.gain_map = gain_map;
    d.offset_map = offset_map;
    d.seed = seed;
    d.frame = frame;
    d.pix_base = pix_base;
    d.frame_dev = frame_dev;
    d.E_out = E_out;
    d.adc_out = adc_out;
    d.adc_f32 = adc_is_float32 ? 1 : 0;
    return splat_run(SMX_SPLAT_ARGS, stages, tile_side, &d);
}

int smx_splat_detect_ts(SMX_SPLAT_PARAMS, const double* gain_map, const double* offset_map,
                        const smx_detector_params_t* params, uint64_t seed, uint32_t frame,
                        uint32_t pix_base, const uint32_t* frame_dev, double* E_out, void* adc_out,
                        int adc_is_float32, int tile_side) {
    return splat_detect_run(SMX_SPLAT_ARGS, gain_map, offset_map, params, seed, frame, pix_base,
                            frame_dev, E_out, adc_out, adc_is_float32, tile_side, 3);
}

int smx_splat_detect_accumulate_ts(SMX_SPLAT_PARAMS, const double* gain_map,
                                   const double* offset_map, const smx_detector_params_t* params,
                                   uint64_t seed, uint32_t frame, uint32_t pix_base,
                                   const uint32_t* frame_dev, double* E_out, void* adc_out,
                                   int adc_is_float32, int tile_side) {
    return splat_detect_run(SMX_SPLAT_ARGS, gain_map, offset_map, params, seed, frame, pix_base,
                            frame_dev, E_out, adc_out, adc_is_float32, tile_side, 2);
}

// LUT mode (smx_lut_t): the same tile machinery on pixel-integral tables per
// depth slice; key[i] is the slice of emitter i.
#define SMX_LUT_PARAMS                                                                           \
    int n, const double *u, const double *v, const double *w, const int *key, const int *fov_id, \
        const smx_fov_t *fovs_dev, int nfov, int n_tiles, const float *lut_tables,               \
        const smx_lut_t *lut, long long list_capacity, int accumulate, double *out,              \
        void *workspace, size_t workspace_bytes, int *status_dev, hipStream_t stream

int smx_splat_lut_ts(SMX_LUT_PARAMS, int stages, int tile_side) {
    if (!lut || stages < 1 || stages > 3) {
        smx::set_error("smx_splat_lut_ts: invalid arguments (stages=%d)", stages);
        return SMX_EVALUE;
    }
    return splat_run(n, u, v, w, key, fov_id, fovs_dev, nfov, n_tiles, nullptr, 0, 0, 1, 2, 1e-9,
                     1.0, list_capacity, accumulate, lut_tables, 1, lut->m, out, workspace,
                     workspace_bytes, status_dev, stream, stages, tile_side, nullptr, lut);
}

int smx_splat_lut_detect_ts(SMX_LUT_PARAMS, const double* gain_map, const double* offset_map,
                            const smx_detector_params_t* params, uint64_t seed, uint32_t frame,
                            uint32_t pix_base, const uint32_t* frame_dev, double* E_out,
                            void* adc_out, int adc_is_float32, int stages, int tile_side) {
    if (!lut || stages < 1 || stages > 3 || !lut_tables || accumulate != 0 || !params ||
        params->emgain <= 0 || !adc_out) {
        smx::set_error("smx_splat_lut_detect_ts: needs tables, accumulate = 0, detector "
                       "parameters with emgain > 0 and an ADC buffer");
        return SMX_EVALUE;
    }
    DetectArgs d;
    d.P = *params;
    d.gain_map = gain_map;
    d.offset_map = offset_map;
    d.seed = seed;
    d.frame = frame;
    d.pix_base = pix_base;
    d.frame_dev = frame_dev;
    d.E_out = E_out;
    d.adc_out = adc_out;
    d.adc_f32 = adc_is_float32 ? 1 : 0;
    return splat_run(n, u, v, w, key, fov_id, fovs_dev, nfov, n_tiles, nullptr, 0, 0, 1, 2, 1e-9,
                     1.0, list_capacity, accumulate, lut_tables, 1, lut->m, out, workspace,
                     workspace_bytes, status_dev, stream, stages, tile_side, &d, lut);
}

}  // extern "C"

// --------------------------------------------------------------------------
// Allocation-free splat: the binning workspace of smx_splat_expected_ts lives
// in `scratch`, a fixed device array of SMX_SPLAT_SCRATCH_BYTES that every
// GPU the library runs on holds until the process ends.  In host code the
// array's name is only a stand-in: its device address on the current GPU
// comes from hipGetSymbolAddress, once per device, and only that address is
// handed to kernels and to the scan.  Every call on one GPU shares that GPU's
// array, so calls that overlap on one GPU (two streams, two threads) race.
// Nothing in the array must start at zero: each call re-initialises the tile
// counters it reads.
// --------------------------------------------------------------------------
#define SMX_SPLAT_SCRATCH_BYTES (64ull << 20)
#define SMX_SCRATCH_MAX_DEVICES 64

static __device__ __align__(256) unsigned char scratch[SMX_SPLAT_SCRATCH_BYTES];

namespace {

std::mutex g_scratch_mu;
void* g_scratch_ptr[SMX_SCRATCH_MAX_DEVICES];

// Device address of `scratch` on the current GPU (looked up on the first call
// for that GPU).
int scratch_address(void** out) {
    int dev = -1;
    if (smx::check_hip(hipGetDevice(&dev), "smx_splat_expected_static: hipGetDevice"))
        return SMX_ERUNTIME;
    if (dev < 0 || dev >= SMX_SCRATCH_MAX_DEVICES) {
        smx::set_error("smx_splat_expected_static: device %d out of range", dev);
        return SMX_ERUNTIME;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            smx::set_error("smx_splat_expected_static: hipGetSymbolAddress(scratch) on device "
                           "%d: %s", dev, hipGetErrorString(e));
            return SMX_ERUNTIME;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return SMX_OK;
}

}  // namespace

extern "C" {

size_t smx_splat_static_scratch_bytes(void) { return SMX_SPLAT_SCRATCH_BYTES; }

int smx_splat_expected_static(int n, const double* u, const double* v, const double* w,
                              const int* key, const int* fov_id, const smx_fov_t* fovs_dev,
                              int nfov, int n_tiles, const double* sat, size_t sat_stride,
                              int n_slots, int ntab, int sat_period, double res, double unit_area,
                              long long list_capacity, int accumulate, const void* K, int k_format,
                              int m, double* out, int* status_dev, hipStream_t stream,
                              int tile_side) {
    if (n < 0 || n_tiles <= 0 || list_capacity < 0) {
        smx::set_error("smx_splat_expected_static: invalid arguments (n=%d, n_tiles=%d)", n,
                       n_tiles);
        return SMX_EVALUE;
    }
    const size_t need = layout_workspace(n, n_tiles, list_capacity, nullptr, nullptr);
    if (need > SMX_SPLAT_SCRATCH_BYTES) {
        smx::set_error("smx_splat_expected_static: %d emitters, %d tiles and %lld list entries "
                       "need %zu bytes of scratch; the fixed array holds %llu (use "
                       "smx_splat_expected_ts with a workspace)",
                       n, n_tiles, list_capacity, need, SMX_SPLAT_SCRATCH_BYTES);
        return SMX_EVALUE;
    }
    void* ws = nullptr;
    const int rc = scratch_address(&ws);
    if (rc != SMX_OK) return rc;
    return splat_run(n, u, v, w, key, fov_id, fovs_dev, nfov, n_tiles, sat, sat_stride, n_slots,
                     ntab, sat_period, res, unit_area, list_capacity, accumulate, K, k_format, m,
                     out, ws, SMX_SPLAT_SCRATCH_BYTES, status_dev, stream, 3, tile_side);
}

__global__ void to_float_kernel(const double* __restrict__ a, size_t n, float* __restrict__ b) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

int smx_phase_table_to_float(const double* K, size_t n, float* K32, hipStream_t stream) {
    if (!K || !K32) {
        smx::set_error("smx_phase_table_to_float: null pointer");
        return SMX_EVALUE;
    }
    if (n == 0) return SMX_OK;
    hipLaunchKernelGGL(to_float_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, K,
                       n, K32);
    SMX_CHECK_LAUNCH("to_float_kernel");
    return SMX_OK;
}


size_t smx_phase_fold_elems(int ntab, int period) {
    if (ntab < 1 || period < 2) return 0;
    return smx::make_phase_fold(ntab, period).elems();
}

int smx_phase_fold_host(int ntab, int period, int ph, int* rank, int* reflected, int* npix) {
    if (ntab < 1 || period < 2 || ph < 0 || ph >= period) {
        smx::set_error("smx_phase_fold_host: invalid arguments (ntab=%d, period=%d, ph=%d)", ntab,
                       period, ph);
        return SMX_EVALUE;
    }
    smx::PhaseFold f = smx::make_phase_fold(ntab, period);
    bool r = false;
    int c = f.canon(ph, r);
    if (rank) *rank = c;
    if (reflected) *reflected = r ? 1 : 0;
    if (npix) *npix = f.npix(ph);
    return SMX_OK;
}

int smx_phase_table_fold(const double* K, int nslots, int ntab, int period, int m, float* Kf,
                         hipStream_t stream) {
    if (!K || !Kf || nslots < 0 || ntab < 1 || period < 2 || m < 1) {
        smx::set_error("smx_phase_table_fold: invalid arguments (period=%d, m=%d)", period, m);
        return SMX_EVALUE;
    }
    smx::PhaseFold f = smx::make_phase_fold(ntab, period);
    if (f.mf > m) {
        smx::set_error("smx_phase_table_fold: block side %d exceeds the table's %d", f.mf, m);
        return SMX_EVALUE;
    }
    if (nslots == 0) return SMX_OK;
    size_t total = (size_t)f.nc * f.nc * f.mf * f.mf * (size_t)nslots;
    hipLaunchKernelGGL(phase_fold_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       stream, K, nslots, period, m, f, Kf);
    SMX_CHECK_LAUNCH("phase_fold_kernel");
    return SMX_OK;
}

}  // extern "C"
