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
