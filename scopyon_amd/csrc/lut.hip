// Radial x depth lookup-table PSF ("LUT mode"): pixel-integral tables per
// depth slice, built in float64 from a float64 Born-Wolf radial table.
//
// Reference: src/scopyon/_epifm.py:98-126 (cartesian kernel: the radial
// profile clamped at r_max on a square of side 2 r_max), _epifm.py:180-213
// (Born-Wolf profile).  The reference integrates each pixel as a sum of 1 nm
// cells whose edges are rounded to the cell grid per emitter
// (_epifm.py:233-282); this mode integrates the same kernel over the exact
// pixel square instead (midpoint rule on cells of pl / (q ns)), tabulated at
// q + 1 sub-pixel phases per axis.
//
// Build per slice: the kernel's values at the cell centres of a fine grid
// covering every pixel any phase can need ((m + 1) pixels per axis), a
// summed-area table of them (float64), and one four-corner difference per
// table entry.  Tables are float32: (q + 1)^2 m^2 entries per slice.
#include "common.hpp"
#include "scopyon_hip_types.h"

namespace {

struct FineGrid {
    int side;          // fine cells per axis: (m + 1) q ns
    int cpp;           // fine cells per pixel: q ns
    double h;          // cell size [m]
    double origin;     // coordinate of cell edge 0 relative to the emitter [m]: -(c + 1) pl
    double r_max;
    double dr;         // radial bin width
    int nb;
};

__device__ inline double kernel_value(const double* __restrict__ prof, const FineGrid& G,
                                      double x, double y) {
    if (fabs(x) > G.r_max || fabs(y) > G.r_max) return 0.0;
    double R = sqrt(x * x + y * y);
    if (R > G.r_max) R = G.r_max;
    const double t = R / G.dr;
    int k = (int)t;
    if (k > G.nb - 2) k = G.nb - 2;
    const double f = t - (double)k;
    return prof[k] + f * (prof[k + 1] - prof[k]);
}

// One wave per fine row: cell values along the row, inclusive scan ->
// sat[z][row + 1][col + 1]; column 0 of the row is 0.
__global__ void __launch_bounds__(256) lut_row_kernel(const double* __restrict__ radial,
                                                      FineGrid G, double* __restrict__ sat) {
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int zi = blockIdx.y;
    if (row >= G.side) return;
    const size_t W = (size_t)G.side + 1;
    const double* prof = radial + (size_t)zi * G.nb;
    double* out = sat + (size_t)zi * W * W + (size_t)(row + 1) * W;
    const double y = G.origin + ((double)row + 0.5) * G.h;
    if (lane == 0) out[0] = 0.0;
    double carry = 0.0;
    for (int c0 = 0; c0 < G.side; c0 += 64) {
        const int col = c0 + lane;
        double v = 0.0;
        if (col < G.side) v = kernel_value(prof, G, G.origin + ((double)col + 0.5) * G.h, y);
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const double o = __shfl_up(v, off, 64);
            if (lane >= off) v += o;
        }
        const double s = carry + v;
        if (col < G.side) out[col + 1] = s;
        carry = __shfl(s, 63, 64);
    }
}

// Running sums down the columns, in place; row 0 of the table is 0.
__global__ void __launch_bounds__(64) lut_col_kernel(FineGrid G, double* __restrict__ sat) {
    const size_t W = (size_t)G.side + 1;
    const int col = blockIdx.x * 64 + threadIdx.x;
    const int zi = blockIdx.y;
    if (col >= (int)W) return;
    double* s = sat + (size_t)zi * W * W;
    s[col] = 0.0;
    double acc = 0.0;
    for (int row = 1; row < (int)W; ++row) {
        acc += s[(size_t)row * W + col];
        s[(size_t)row * W + col] = acc;
    }
}

// K[z][qu][qv][du][dv] = h^2 * (cells of pixel (du, dv) at phases (qu, qv)).
__global__ void lut_phase_kernel(const double* __restrict__ sat, FineGrid G, int q, int ns, int m,
                                 int nz, float* __restrict__ K) {
    const size_t per = (size_t)(q + 1) * (q + 1) * m * m;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= per * nz) return;
    size_t r = idx;
    const int dv = (int)(r % m); r /= m;
    const int du = (int)(r % m); r /= m;
    const int qv = (int)(r % (q + 1)); r /= (q + 1);
    const int qu = (int)(r % (q + 1));
    const int zi = (int)(r / (q + 1));
    const size_t W = (size_t)G.side + 1;
    const double* S = sat + (size_t)zi * W * W;
    const int u0 = (du + 1) * G.cpp - qu * ns, u1 = u0 + G.cpp;
    const int v0 = (dv + 1) * G.cpp - qv * ns, v1 = v0 + G.cpp;
    const double sum = ((S[(size_t)u1 * W + v1] - S[(size_t)u0 * W + v1]) - S[(size_t)u1 * W + v0]) +
                       S[(size_t)u0 * W + v0];
    K[idx] = (float)(sum * (G.h * G.h));
}

// Cell-interleaved copy for the splat: cell (cu, cv), cu, cv < q, holds per
// footprint pixel the four phases around it, in the order of the bilinear
// taps: (cu, cv), (cu, cv + 1), (cu + 1, cv), (cu + 1, cv + 1).  One emitter
// then reads one contiguous 16-byte entry per pixel.
__global__ void lut_cells_kernel(const float* __restrict__ K, int q, int m, int nz,
                                 float4* __restrict__ K4) {
    const size_t mm = (size_t)m * m;
    const size_t per = (size_t)q * q * mm;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= per * nz) return;
    const size_t pix = idx % mm;
    size_t r = idx / mm;
    const int cv = (int)(r % q); r /= q;
    const int cu = (int)(r % q);
    const size_t zi = r / q;
    const float* Kz = K + zi * (size_t)(q + 1) * (q + 1) * mm;
    auto at = [&](int a, int b) { return Kz[((size_t)a * (q + 1) + b) * mm + pix]; };
    K4[idx] = make_float4(at(cu, cv), at(cu, cv + 1), at(cu + 1, cv), at(cu + 1, cv + 1));
}

bool lut_ok(const smx_lut_t* L) {
    return L && L->q >= 1 && L->q <= 256 && L->ns >= 1 && L->c >= 1 && L->m == 2 * L->c + 1 &&
           L->nslices >= 1 && L->pixel_length > 0.0 && L->r_max > 0.0 &&
           (double)L->c >= L->r_max / L->pixel_length &&
           (size_t)(L->m + 1) * L->q * L->ns <= 65536;
}

FineGrid fine_grid(const smx_lut_t* L, int nb) {
    FineGrid G;
    G.cpp = L->q * L->ns;
    G.side = (L->m + 1) * G.cpp;
    G.h = L->pixel_length / (double)G.cpp;
    G.origin = -(double)(L->c + 1) * L->pixel_length;
    G.r_max = L->r_max;
    G.nb = nb;
    G.dr = L->r_max / (double)(nb - 1);
    return G;
}

}  // namespace

extern "C" {

size_t smx_lut_table_elems(const smx_lut_t* lut) {
    if (!lut_ok(lut)) return 0;
    return (size_t)(lut->q + 1) * (lut->q + 1) * lut->m * lut->m;
}

size_t smx_lut_cell_elems(const smx_lut_t* lut) {
    if (!lut_ok(lut)) return 0;
    return (size_t)lut->q * lut->q * lut->m * lut->m * 4;
}

int smx_lut_cells(const float* K, int nz, const smx_lut_t* lut, float* cells, hipStream_t stream) {
    if (!lut_ok(lut) || nz < 0 || !K || !cells) {
        smx::set_error("smx_lut_cells: invalid arguments (nz=%d)", nz);
        return SMX_EVALUE;
    }
    const size_t total = smx_lut_cell_elems(lut) / 4 * (size_t)nz;
    if (total == 0) return SMX_OK;
    hipLaunchKernelGGL(lut_cells_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       stream, K, lut->q, lut->m, nz, (float4*)cells);
    SMX_CHECK_LAUNCH("lut_cells_kernel");
    return SMX_OK;
}

size_t smx_lut_build_workspace_bytes(const smx_lut_t* lut, int nz_batch) {
    if (!lut_ok(lut) || nz_batch < 1) return 0;
    const size_t W = (size_t)(lut->m + 1) * lut->q * lut->ns + 1;
    return W * W * sizeof(double) * (size_t)nz_batch;
}

int smx_lut_build(const double* radial, int nz, int nb, const smx_lut_t* lut, double* workspace,
                  size_t workspace_bytes, float* K, hipStream_t stream) {
    if (!lut_ok(lut) || nz < 0 || nb < 2 || !radial || !workspace || !K) {
        smx::set_error("smx_lut_build: invalid arguments (nz=%d, nb=%d)", nz, nb);
        return SMX_EVALUE;
    }
    const size_t per_slice = smx_lut_build_workspace_bytes(lut, 1);
    const int batch = (int)(workspace_bytes / per_slice);
    if (batch < 1) {
        smx::set_error("smx_lut_build: workspace too small (%zu < %zu bytes)", workspace_bytes,
                       per_slice);
        return SMX_EVALUE;
    }
    const FineGrid G = fine_grid(lut, nb);
    const size_t elems = smx_lut_table_elems(lut);
    for (int z0 = 0; z0 < nz; z0 += batch) {
        const int nzb = nz - z0 < batch ? nz - z0 : batch;
        hipLaunchKernelGGL(lut_row_kernel, dim3((G.side + 3) / 4, nzb), dim3(256), 0, stream,
                           radial + (size_t)z0 * nb, G, workspace);
        SMX_CHECK_LAUNCH("lut_row_kernel");
        hipLaunchKernelGGL(lut_col_kernel, dim3((G.side + 1 + 63) / 64, nzb), dim3(64), 0, stream, G,
                           workspace);
        SMX_CHECK_LAUNCH("lut_col_kernel");
        const size_t total = elems * nzb;
        hipLaunchKernelGGL(lut_phase_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                           stream, workspace, G, lut->q, lut->ns, lut->m, nzb,
                           K + (size_t)z0 * elems);
        SMX_CHECK_LAUNCH("lut_phase_kernel");
    }
    return SMX_OK;
}

}  // extern "C"
