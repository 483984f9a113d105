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
// Work decomposition: the frame is cut into 16x16-pixel tiles; emitters are
// binned to the tiles their footprint touches (count -> scan -> fill), each
// tile's list is sorted by emitter index in LDS (so every pixel adds its
// contributions in emitter order, like the reference loop, and the result is
// run-to-run deterministic), then one workgroup per tile accumulates its 256
// pixels in registers.  No float atomics.
#include "common.hpp"
#include "scopyon_hip_types.h"

#include <hipcub/hipcub.hpp>

namespace {

constexpr int TS = 16;         // tile side in pixels
constexpr int TB = TS + 1;     // boundaries per tile axis
constexpr int CHUNK = 8;       // emitters staged in LDS per round
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
                                 int n_tiles, int* __restrict__ tile_count) {
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
            if (tile >= 0 && tile < n_tiles) atomicAdd(&tile_count[tile], 1);
        }
}

__global__ void bin_fill_kernel(int n, const double* __restrict__ u, const double* __restrict__ v,
                                const double* __restrict__ w, const int* __restrict__ fov_id,
                                const smx_fov_t* __restrict__ fovs, int nfov, double ext_half,
                                int n_tiles, const int* __restrict__ tile_begin,
                                int* __restrict__ tile_cursor,
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
            int pos = atomicAdd(&tile_cursor[tile], 1);
            long long at = (long long)tile_begin[tile] + pos;
            if (at < capacity) list[at] = i;
            else atomicOr(status, 2);
        }
}

__global__ void fill_int_kernel(int* __restrict__ p, size_t n, int value) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
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
    const int* list;
    long long capacity;      // entries in list
    const double* sat;
    size_t sat_stride;
    int n_slots;             // tables available in sat
    smx::SatLayout L;
    double res, ext_half, unit_area;
    int ntab;
    int accumulate;          // 1: out += acc ; 0: out = acc
    double* out;
    int* status;             // set to 1 if a tile list had to be processed unsorted
};

__device__ void lds_bitonic_sort(int* a, int n_pow2) {
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

__global__ void __launch_bounds__(256) splat_tile_kernel(TileArgs A) {
    __shared__ int s_list[SORT_CAP];
    __shared__ int s_rowB[CHUNK][TB];
    __shared__ int s_colB[CHUNK][TB];
    __shared__ double s_corner[CHUNK][TB][TB];
    __shared__ double s_w[CHUNK];
    __shared__ long long s_sat_off[CHUNK];

    const int tile = blockIdx.x;
    const int fov = A.tile_fov[tile];
    if (fov < 0 || fov >= A.nfov) return;   // tile not covered by any FOV
    const smx_fov_t F = A.fovs[fov];
    const int local = tile - F.tile_off;
    const int tr = local / F.tiles_c, tc = local % F.tiles_c;
    const int ty = threadIdx.x / TS, tx = threadIdx.x % TS;
    const int pi = tr * TS + ty, pj = tc * TS + tx;   // this thread's pixel (row, col)
    const bool inside = pi < F.nw && pj < F.nh;

    const int begin = A.tile_begin[tile];
    int count = A.tile_begin[tile + 1] - begin;
    if ((long long)begin + count > A.capacity) count = (int)max(0LL, A.capacity - begin);
    const int* list = A.list + begin;

    // sort the tile list by emitter index (emitter order per pixel)
    bool staged = count <= SORT_CAP;
    if (staged && count > 0) {
        int np2 = 1;
        while (np2 < count) np2 <<= 1;
        for (int k = threadIdx.x; k < np2; k += blockDim.x) s_list[k] = k < count ? list[k] : 0x7fffffff;
        __syncthreads();
        lds_bitonic_sort(s_list, np2);
    } else if (!staged && threadIdx.x == 0) {
        atomicOr(A.status, 1);
    }

    double acc = 0.0;
    for (int base = 0; base < count; base += CHUNK) {
        const int nc = min(CHUNK, count - base);
        // phase A: row/column cell boundaries of this tile for nc emitters
        for (int t = threadIdx.x; t < nc * 2 * TB; t += blockDim.x) {
            int c = t / (2 * TB), m = t % (2 * TB);
            int e = staged ? s_list[base + c] : list[base + c];
            bool is_row = m < TB;
            int mm = is_row ? m : m - TB;
            double xi = is_row ? A.u[e] - F.p0u : A.v[e] - F.p0v;
            double half = is_row ? F.half_w : F.half_h;
            int npix = is_row ? F.nw : F.nh;
            Axis ax = make_axis(half, xi, A.ext_half, F.pixel_length, npix);
            long long k = (long long)(is_row ? tr : tc) * TS + mm;
            int b = -1;
            if (k >= ax.k0 && k <= ax.k1) b = axis_boundary(ax, k, F.pixel_length, A.res, A.ntab);
            if (is_row) s_rowB[c][mm] = b; else s_colB[c][mm] = b;
            if (m == 0) {
                int slot = A.key[e];
                bool ok = slot >= 0 && slot < A.n_slots;
                s_w[c] = ok ? A.w[e] : 0.0;
                s_sat_off[c] = ok ? (long long)slot * (long long)A.sat_stride : -1;
                if (!ok) atomicOr(A.status, 8);
            }
        }
        __syncthreads();
        // phase B: gather the SAT corners this tile needs
        for (int t = threadIdx.x; t < nc * TB * TB; t += blockDim.x) {
            int c = t / (TB * TB), r = t % (TB * TB);
            int m = r / TB, l = r % TB;
            int b = s_rowB[c][m], cc = s_colB[c][l];
            double val = 0.0;
            if (b >= 0 && cc >= 0 && s_sat_off[c] >= 0) val = A.sat[s_sat_off[c] + A.L.index(b, cc)];
            s_corner[c][m][l] = val;
        }
        __syncthreads();
        // phase C: per-pixel accumulation in emitter order
        if (inside) {
            for (int c = 0; c < nc; ++c) {
                int b0 = s_rowB[c][ty], b1 = s_rowB[c][ty + 1];
                int c0 = s_colB[c][tx], c1 = s_colB[c][tx + 1];
                if (b0 < 0 || b1 < 0 || c0 < 0 || c1 < 0 || s_sat_off[c] < 0) continue;
                if (b1 <= b0 || c1 <= c0) continue;
                double s = ((s_corner[c][ty + 1][tx + 1] - s_corner[c][ty][tx + 1])
                            - s_corner[c][ty + 1][tx]) + s_corner[c][ty][tx];
                double photons = s * A.unit_area;
                if (photons > 0.0) acc += photons * s_w[c];
            }
        }
        __syncthreads();
    }
    if (inside) {
        size_t o = (size_t)F.pix_off + (size_t)pi * F.nh + pj;
        if (A.accumulate) {
            if (count > 0) A.out[o] = A.out[o] + acc;
        } else {
            A.out[o] = acc;
        }
    }
}

__global__ void tile_fov_kernel(int nfov, const smx_fov_t* __restrict__ fovs, int* __restrict__ tile_fov,
                                int n_tiles) {
    int f = blockIdx.x;
    if (f >= nfov) return;
    const smx_fov_t F = fovs[f];
    int nt = F.tiles_r * F.tiles_c;
    for (int t = threadIdx.x; t < nt; t += blockDim.x) {
        int g = F.tile_off + t;
        if (g >= 0 && g < n_tiles) tile_fov[g] = f;
    }
}

struct SplatWorkspace {
    int* tile_count;   // n_tiles + 1
    int* tile_begin;   // n_tiles + 1
    int* tile_cursor;  // n_tiles
    int* tile_fov;     // n_tiles
    int* list;         // capacity
    void* cub_tmp;
    size_t cub_bytes;
};

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

size_t scan_tmp_bytes(int n_tiles) {
    size_t bytes = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (int*)nullptr, (int*)nullptr, n_tiles + 1);
    return bytes;
}

size_t layout_workspace(int n_tiles, long long capacity, char* base, SplatWorkspace* ws) {
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += align256(bytes); return o; };
    size_t o_count = take(sizeof(int) * (n_tiles + 1));
    size_t o_begin = take(sizeof(int) * (n_tiles + 1));
    size_t o_cursor = take(sizeof(int) * n_tiles);
    size_t o_fov = take(sizeof(int) * n_tiles);
    size_t o_list = take(sizeof(int) * (size_t)(capacity > 0 ? capacity : 1));
    size_t cb = scan_tmp_bytes(n_tiles);
    size_t o_cub = take(cb);
    if (ws && base) {
        ws->tile_count = (int*)(base + o_count);
        ws->tile_begin = (int*)(base + o_begin);
        ws->tile_cursor = (int*)(base + o_cursor);
        ws->tile_fov = (int*)(base + o_fov);
        ws->list = (int*)(base + o_list);
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

size_t smx_splat_workspace_bytes(int n_tiles, long long list_capacity) {
    if (n_tiles <= 0 || list_capacity < 0) return 0;
    return layout_workspace(n_tiles, list_capacity, nullptr, nullptr);
}

int smx_splat_expected(int n, const double* u, const double* v, const double* w, const int* key,
                       const int* fov_id, const smx_fov_t* fovs_dev, int nfov, int n_tiles,
                       const double* sat, size_t sat_stride, int n_slots, int ntab, int sat_period,
                       double res, double unit_area, long long list_capacity, int accumulate,
                       double* out, void* workspace, size_t workspace_bytes, int* status_dev,
                       hipStream_t stream) {
    if (n < 0 || nfov <= 0 || n_tiles <= 0 || ntab < 1 || sat_period < 1 || !fovs_dev || !out ||
        !workspace || !status_dev || (n > 0 && (!u || !v || !w || !key || !sat))) {
        smx::set_error("smx_splat_expected: invalid arguments (n=%d, nfov=%d, n_tiles=%d)", n,
                       nfov, n_tiles);
        return SMX_EVALUE;
    }
    size_t need = layout_workspace(n_tiles, list_capacity, nullptr, nullptr);
    if (workspace_bytes < need) {
        smx::set_error("smx_splat_expected: workspace too small (%zu < %zu bytes)",
                       workspace_bytes, need);
        return SMX_EVALUE;
    }
    SplatWorkspace ws;
    layout_workspace(n_tiles, list_capacity, (char*)workspace, &ws);
    const double ext_half = (res * (double)(ntab - 1)) * 0.5;   // signal_width * 0.5
    const int nb = (n + 255) / 256;

    hipLaunchKernelGGL(fill_int_kernel, dim3((n_tiles + 1 + 255) / 256), dim3(256), 0, stream,
                       ws.tile_count, (size_t)n_tiles + 1, 0);
    SMX_CHECK_LAUNCH("fill_int_kernel");
    hipLaunchKernelGGL(fill_int_kernel, dim3((n_tiles + 255) / 256), dim3(256), 0, stream,
                       ws.tile_cursor, (size_t)n_tiles, 0);
    SMX_CHECK_LAUNCH("fill_int_kernel");
    hipLaunchKernelGGL(fill_int_kernel, dim3((n_tiles + 255) / 256), dim3(256), 0, stream,
                       ws.tile_fov, (size_t)n_tiles, -1);
    SMX_CHECK_LAUNCH("fill_int_kernel");
    hipLaunchKernelGGL(tile_fov_kernel, dim3(nfov), dim3(256), 0, stream, nfov, fovs_dev,
                       ws.tile_fov, n_tiles);
    SMX_CHECK_LAUNCH("tile_fov_kernel");
    if (n > 0) {
        hipLaunchKernelGGL(bin_count_kernel, dim3(nb), dim3(256), 0, stream, n, u, v, w, fov_id,
                           fovs_dev, nfov, ext_half, n_tiles, ws.tile_count);
        SMX_CHECK_LAUNCH("bin_count_kernel");
    }
    size_t cb = ws.cub_bytes;
    if (smx::check_hip(hipcub::DeviceScan::ExclusiveSum(ws.cub_tmp, cb, ws.tile_count,
                                                        ws.tile_begin, n_tiles + 1, stream),
                       "hipcub ExclusiveSum"))
        return SMX_ERUNTIME;
    if (n > 0) {
        hipLaunchKernelGGL(bin_fill_kernel, dim3(nb), dim3(256), 0, stream, n, u, v, w, fov_id,
                           fovs_dev, nfov, ext_half, n_tiles, ws.tile_begin, ws.tile_cursor, ws.list,
                           list_capacity, status_dev);
        SMX_CHECK_LAUNCH("bin_fill_kernel");
    }
    TileArgs A;
    A.u = u; A.v = v; A.w = w; A.key = key; A.fov_id = fov_id; A.fovs = fovs_dev; A.nfov = nfov;
    A.tile_fov = ws.tile_fov; A.tile_begin = ws.tile_begin; A.list = ws.list;
    A.capacity = list_capacity;
    A.sat = sat; A.sat_stride = sat_stride; A.n_slots = n_slots; A.L = smx::make_sat_layout(ntab + 1, sat_period);
    A.res = res; A.ext_half = ext_half; A.unit_area = unit_area; A.ntab = ntab;
    A.accumulate = accumulate; A.out = out; A.status = status_dev;
    hipLaunchKernelGGL(splat_tile_kernel, dim3(n_tiles), dim3(256), 0, stream, A);
    SMX_CHECK_LAUNCH("splat_tile_kernel");
    return SMX_OK;
}

}  // extern "C"
