// Brownian particle sampling.
//
// Reference: src/scopyon/sampling2.py:33-68 (per-state D, periodic wrap,
//            state transitions), src/scopyon/sampling.py:88-119 (per-axis D).
//
//  * predrawn : replays a whole trajectory from normals/uniforms drawn by the
//               caller in the reference's RandomState order; bit-exact.
//               One thread per particle walks all time steps.
//  * philox   : GPU-native step on structure-of-arrays coordinates with
//               per-particle D, Philox4x32-10 normals and a boundary mode
//               (none / periodic / reflecting).
#include "common.hpp"

namespace {

// Reflecting walls on [lo, hi]: fold x into [lo, lo + 2L) and mirror the
// upper half (exact image method for any number of wall crossings).
__device__ inline double reflect(double x, double lo, double hi) {
    double L = hi - lo;
    if (!(L > 0.0)) return lo;
    double m = smx::np_remainder(x - lo, 2.0 * L);
    return lo + (L - fabs(m - L));
}

__global__ void brownian_predrawn_kernel(int n, int ndim, int nsteps, int nstates,
                                         const double* __restrict__ x0,
                                         const int* __restrict__ state0,
                                         const double* __restrict__ scale,
                                         const double* __restrict__ xi, int boundary,
                                         const double* __restrict__ lower,
                                         const double* __restrict__ upper,
                                         const double* __restrict__ pacc,
                                         const double* __restrict__ trans_u,
                                         double* __restrict__ xs, int* __restrict__ states) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x[8];
    for (int d = 0; d < ndim; ++d) {
        x[d] = x0[(size_t)i * ndim + d];
        xs[(size_t)i * ndim + d] = x[d];
    }
    int s = state0 ? state0[i] : 0;
    if (states) states[i] = s;
    s = s < 0 ? 0 : (s >= nstates ? nstates - 1 : s);   // keep table reads in bounds
    for (int t = 0; t < nsteps; ++t) {
        const double* sc = scale + ((size_t)t * nstates + s) * ndim;
        const double* z = xi + ((size_t)t * n + i) * ndim;
        for (int d = 0; d < ndim; ++d) x[d] = x[d] + (0.0 + sc[d] * z[d]);
        if (boundary == 1) {
            for (int d = 0; d < ndim; ++d) {
                if (upper[d] > lower[d])
                    x[d] = smx::np_remainder(x[d] - lower[d], upper[d] - lower[d]) + lower[d];
                else
                    x[d] = lower[d];
            }
        } else if (boundary == 2) {
            for (int d = 0; d < ndim; ++d) x[d] = reflect(x[d], lower[d], upper[d]);
        }
        if (pacc) {
            // searchsorted(Pacc[state], u, side='left'): count of entries < u
            double u = trans_u[(size_t)t * n + i];
            const double* row = pacc + ((size_t)t * nstates + s) * nstates;
            int k = 0;
            while (k < nstates && row[k] < u) ++k;
            s = k < nstates ? k : nstates - 1;
        }
        double* out = xs + ((size_t)(t + 1) * n + i) * ndim;
        for (int d = 0; d < ndim; ++d) out[d] = x[d];
        if (states) states[(size_t)(t + 1) * n + i] = s;
    }
}

// Box-Muller normals of the Philox paths in float32 from 53-bit uniforms:
// u1 in (0, 1] converted to float still reaches 2^-53, so the tails reach
// |z| ~ 8.6 as in float64, and the ~1e-7 relative rounding of a displacement
// is far below anything physical; float64 log / sqrt / sincos took most of a
// step's time (640,000 emitters: ~10 us).
__device__ inline void normal_pair(const smx::u32x4& r, double& z0, double& z1) {
    const float u1 = (float)smx::u01_53_open0(r.x, r.y);
    const float u2 = (float)smx::u01_53(r.z, r.w);
    const float rad = sqrtf(-2.0f * logf(u1));
    float s, c;
    sincospif(2.0f * u2, &s, &c);
    z0 = (double)(rad * c);
    z1 = (double)(rad * s);
}
__device__ inline double normal_one(const smx::u32x4& r) {
    const float u1 = (float)smx::u01_53_open0(r.x, r.y);
    const float u2 = (float)smx::u01_53(r.z, r.w);
    return (double)(sqrtf(-2.0f * logf(u1)) * cospif(2.0f * u2));
}

// One step for structure-of-arrays coordinates (ndim <= 3).
__global__ void brownian_philox_kernel(int n, int ndim, double* __restrict__ x0,
                                       double* __restrict__ x1, double* __restrict__ x2,
                                       const double* __restrict__ D, double D_scalar, double dt,
                                       int boundary, double lo0, double lo1, double lo2,
                                       double hi0, double hi1, double hi2, uint64_t seed,
                                       uint32_t step, uint32_t id_base,
                                       const int* __restrict__ fov_id, int nfov,
                                       const double* __restrict__ fov_box,
                                       const uint32_t* __restrict__ step_dev) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (step_dev) step += *step_dev;
    double Di = D ? D[i] : D_scalar;
    double sc = sqrt(2.0 * Di * dt);
    smx::u32x4 r = smx::philox4x32_10_dev(smx::u32x4{id_base + (uint32_t)i, step, 0xB0A1u, 0u},
                                      (uint32_t)seed, (uint32_t)(seed >> 32));
    double z[3] = {0.0, 0.0, 0.0};
    normal_pair(r, z[0], z[1]);
    if (ndim > 2) {
        smx::u32x4 q = smx::philox4x32_10_dev(smx::u32x4{id_base + (uint32_t)i, step, 0xB0A1u, 1u},
                                          (uint32_t)seed, (uint32_t)(seed >> 32));
        z[2] = normal_one(q);
    }
    double* xs[3] = {x0, x1, x2};
    double lo[3] = {lo0, lo1, lo2}, hi[3] = {hi0, hi1, hi2};
    if (fov_box) {
        // per-FOV walls: fov_box[f] = {lo0, lo1, lo2, hi0, hi1, hi2}
        int f = fov_id ? fov_id[i] : 0;
        f = f < 0 ? 0 : (f >= nfov ? nfov - 1 : f);
        const double* b = fov_box + (size_t)f * 6;
        for (int d = 0; d < 3; ++d) { lo[d] = b[d]; hi[d] = b[3 + d]; }
    }
    for (int d = 0; d < ndim; ++d) {
        double v = xs[d][i] + sc * z[d];
        if (boundary == 1) {
            v = (hi[d] > lo[d]) ? smx::np_remainder(v - lo[d], hi[d] - lo[d]) + lo[d] : lo[d];
        } else if (boundary == 2) {
            v = reflect(v, lo[d], hi[d]);
        }
        xs[d][i] = v;
    }
}

// One step with per-state, per-axis diffusion and optional state
// transitions (sampling2.py:33-68 with a per-axis D as in sampling.py:115-118):
// x_d += sqrt(2 D[s][d] dt) N(0, 1) from the same Philox blocks as
// brownian_philox_kernel, the boundary, then s = searchsorted(pacc[s], u,
// 'left') with u from block 2 of the particle's stream (transitions after the
// move, as the reference orders them).
__global__ void brownian_states_kernel(int n, int ndim, double* __restrict__ x0,
                                       double* __restrict__ x1, double* __restrict__ x2,
                                       int* __restrict__ state, int nstates,
                                       const double* __restrict__ D_tab, double dt, int boundary,
                                       double lo0, double lo1, double lo2, double hi0, double hi1,
                                       double hi2, const double* __restrict__ pacc, uint64_t seed,
                                       uint32_t step, uint32_t id_base) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int s = state ? state[i] : 0;
    s = s < 0 ? 0 : (s >= nstates ? nstates - 1 : s);
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    smx::u32x4 r = smx::philox4x32_10_dev(smx::u32x4{id_base + (uint32_t)i, step, 0xB0A1u, 0u}, k0, k1);
    double z[3] = {0.0, 0.0, 0.0};
    normal_pair(r, z[0], z[1]);
    if (ndim > 2) {
        smx::u32x4 q = smx::philox4x32_10_dev(smx::u32x4{id_base + (uint32_t)i, step, 0xB0A1u, 1u}, k0, k1);
        z[2] = normal_one(q);
    }
    double* xs[3] = {x0, x1, x2};
    const double lo[3] = {lo0, lo1, lo2}, hi[3] = {hi0, hi1, hi2};
    for (int d = 0; d < ndim; ++d) {
        double v = xs[d][i] + sqrt(2.0 * D_tab[(size_t)s * ndim + d] * dt) * z[d];
        if (boundary == 1) {
            v = (hi[d] > lo[d]) ? smx::np_remainder(v - lo[d], hi[d] - lo[d]) + lo[d] : lo[d];
        } else if (boundary == 2) {
            v = reflect(v, lo[d], hi[d]);
        }
        xs[d][i] = v;
    }
    if (pacc && state) {
        smx::u32x4 t = smx::philox4x32_10_dev(smx::u32x4{id_base + (uint32_t)i, step, 0xB0A1u, 2u}, k0, k1);
        const double u = smx::u01_53(t.x, t.y);
        const double* row = pacc + (size_t)s * nstates;
        int k = 0;
        while (k < nstates && row[k] < u) ++k;
        state[i] = k < nstates ? k : nstates - 1;
    }
}

__global__ void uniform_init_kernel(int n, int ndim, double* __restrict__ x0,
                                    double* __restrict__ x1, double* __restrict__ x2,
                                    double lo0, double lo1, double lo2, double hi0, double hi1,
                                    double hi2, uint64_t seed, uint32_t seq, uint32_t id_base) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    smx::u32x4 r = smx::philox4x32_10_dev(smx::u32x4{id_base + (uint32_t)i, seq, 0x1217u, 0u},
                                      (uint32_t)seed, (uint32_t)(seed >> 32));
    smx::u32x4 q = smx::philox4x32_10_dev(smx::u32x4{id_base + (uint32_t)i, seq, 0x1217u, 1u},
                                      (uint32_t)seed, (uint32_t)(seed >> 32));
    double u[3] = {smx::u01_53(r.x, r.y), smx::u01_53(r.z, r.w), smx::u01_53(q.x, q.y)};
    double* xs[3] = {x0, x1, x2};
    double lo[3] = {lo0, lo1, lo2}, hi[3] = {hi0, hi1, hi2};
    for (int d = 0; d < ndim; ++d)
        xs[d][i] = lo[d] < hi[d] ? lo[d] + (hi[d] - lo[d]) * u[d] : lo[d];
}

}  // namespace

extern "C" {

int smx_brownian_predrawn(int n, int ndim, int nsteps, int nstates, const double* x0,
                          const int* state0, const double* scale, const double* xi, int boundary,
                          const double* lower, const double* upper, const double* pacc,
                          const double* trans_u, double* xs, int* states, hipStream_t stream) {
    if (n < 0 || ndim < 1 || ndim > 8 || nsteps < 0 || nstates < 1 || boundary < 0 ||
        boundary > 2 ||
        (n > 0 && (!x0 || !xs || (nsteps > 0 && (!scale || !xi)) ||
                   (boundary && (!lower || !upper)) || (pacc && !trans_u)))) {
        smx::set_error("smx_brownian_predrawn: invalid arguments (n=%d, ndim=%d, nsteps=%d)", n,
                       ndim, nsteps);
        return SMX_EVALUE;
    }
    if (n == 0) return SMX_OK;
    hipLaunchKernelGGL(brownian_predrawn_kernel, dim3((n + 127) / 128), dim3(128), 0, stream, n,
                       ndim, nsteps, nstates, x0, state0, scale, xi, boundary, lower, upper, pacc,
                       trans_u, xs, states);
    SMX_CHECK_LAUNCH("brownian_predrawn_kernel");
    return SMX_OK;
}

int smx_brownian_philox(int n, int ndim, double* x0, double* x1, double* x2, const double* D,
                        double D_scalar, double dt, int boundary, const double* lower,
                        const double* upper, uint64_t seed, uint32_t step, uint32_t id_base,
                        const int* fov_id, int nfov, const double* fov_box,
                        const uint32_t* step_dev, hipStream_t stream) {
    if (n < 0 || ndim < 1 || ndim > 3 || boundary < 0 || boundary > 2 ||
        (n > 0 && (!x0 || (ndim > 1 && !x1) || (ndim > 2 && !x2))) ||
        (boundary && !fov_box && (!lower || !upper)) || (fov_box && nfov < 1)) {
        smx::set_error("smx_brownian_philox: invalid arguments (n=%d, ndim=%d, boundary=%d)", n,
                       ndim, boundary);
        return SMX_EVALUE;
    }
    if (n == 0) return SMX_OK;
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    if (boundary && !fov_box)
        for (int d = 0; d < ndim; ++d) { lo[d] = lower[d]; hi[d] = upper[d]; }
    hipLaunchKernelGGL(brownian_philox_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, n,
                       ndim, x0, x1, x2, D, D_scalar, dt, boundary, lo[0], lo[1], lo[2], hi[0],
                       hi[1], hi[2], seed, step, id_base, fov_id, nfov, fov_box, step_dev);
    SMX_CHECK_LAUNCH("brownian_philox_kernel");
    return SMX_OK;
}

int smx_brownian_philox_states(int n, int ndim, double* x0, double* x1, double* x2, int* state,
                               int nstates, const double* D_tab, double dt, int boundary,
                               const double* lower, const double* upper, const double* pacc,
                               uint64_t seed, uint32_t step, uint32_t id_base,
                               hipStream_t stream) {
    if (n < 0 || ndim < 1 || ndim > 3 || nstates < 1 || boundary < 0 || boundary > 2 || !D_tab ||
        (n > 0 && (!x0 || (ndim > 1 && !x1) || (ndim > 2 && !x2))) ||
        (boundary && (!lower || !upper)) || (pacc && !state)) {
        smx::set_error("smx_brownian_philox_states: invalid arguments (n=%d, ndim=%d, "
                       "nstates=%d)", n, ndim, nstates);
        return SMX_EVALUE;
    }
    if (n == 0) return SMX_OK;
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    if (boundary)
        for (int d = 0; d < ndim; ++d) { lo[d] = lower[d]; hi[d] = upper[d]; }
    hipLaunchKernelGGL(brownian_states_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, n,
                       ndim, x0, x1, x2, state, nstates, D_tab, dt, boundary, lo[0], lo[1], lo[2],
                       hi[0], hi[1], hi[2], pacc, seed, step, id_base);
    SMX_CHECK_LAUNCH("brownian_states_kernel");
    return SMX_OK;
}

int smx_uniform_init(int n, int ndim, double* x0, double* x1, double* x2, const double* lower,
                     const double* upper, uint64_t seed, uint32_t seq, uint32_t id_base,
                     hipStream_t stream) {
    if (n < 0 || ndim < 1 || ndim > 3 || !lower || !upper ||
        (n > 0 && (!x0 || (ndim > 1 && !x1) || (ndim > 2 && !x2)))) {
        smx::set_error("smx_uniform_init: invalid arguments (n=%d, ndim=%d)", n, ndim);
        return SMX_EVALUE;
    }
    if (n == 0) return SMX_OK;
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int d = 0; d < ndim; ++d) { lo[d] = lower[d]; hi[d] = upper[d]; }
    hipLaunchKernelGGL(uniform_init_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, n, ndim,
                       x0, x1, x2, lo[0], lo[1], lo[2], hi[0], hi[1], hi[2], seed, seq, id_base);
    SMX_CHECK_LAUNCH("uniform_init_kernel");
    return SMX_OK;
}

__global__ void counter_add_kernel(uint32_t* c, uint32_t by) {
    if (threadIdx.x == 0 && blockIdx.x == 0) c[0] += by;
}

int smx_counter_add(uint32_t* counter, uint32_t by, hipStream_t stream) {
    if (!counter) {
        smx::set_error("smx_counter_add: null counter");
        return SMX_EVALUE;
    }
    hipLaunchKernelGGL(counter_add_kernel, dim3(1), dim3(64), 0, stream, counter, by);
    SMX_CHECK_LAUNCH("counter_add_kernel");
    return SMX_OK;
}

}  // extern "C"
