// Philox EMCCD sampler (float32), shared by the detector kernel
// (emccd_queue_kernel, detector.hip) and the fused splat + detector tile pass
// (splat_phase_kernel with its detector epilogue, splat.hip): both run this
// code, so a pixel's draws do not depend on which kernel makes them.
//
// Reference: src/scopyon/_epifm.py:1448-1454 (noise drawn, then the EM
// register), _epifm.py:355-416 (EMCCD noise and register).
#ifndef SMX_EMCCD_BLOCK_HPP
#define SMX_EMCCD_BLOCK_HPP

#include "common.hpp"
#include "scopyon_hip_types.h"

namespace smx {
namespace emccd {

// ---------------------------------------------------------------------------
// Philox samplers (float32 arithmetic: these paths are statistical, not
// bit-parity paths).  Every draw comes from a fixed Philox block of the
// pixel's stream -- block 0: readout noise + Poisson uniform + one spare
// word; blocks 1..: PTRS iterations; blocks 2^16..: Gamma iterations -- so
// the sampler needs no stream state and stays inlined (no call spills).
// ---------------------------------------------------------------------------
__device__ inline float u01f(uint32_t w) { return (float)(w >> 8) * 0x1.0p-24f; }   // [0, 1)
// natural log on the hardware log2 (v_log_f32, ~1e-7 relative)
__device__ inline float ln_hw(float v) { return 0.69314718055994531f * __builtin_amdgcn_logf(v); }
__device__ inline float u01f_open0(uint32_t w) { return 1.0f - u01f(w); }          // (0, 1]

struct PixelRng {
    uint32_t id, frame, k0, k1;
    __device__ smx::u32x4 block(uint32_t b) const {
        return smx::philox4x32_10_dev(smx::u32x4{id, frame, 0x0E3CCD01u, b}, k0, k1);
    }
};

// Box-Muller on the hardware transcendentals: v_log_f32 is log2, v_cos_f32
// takes its argument in revolutions (cos(2 pi v)); both ~1e-7 accurate on
// these ranges.
__device__ inline float normal_f(uint32_t a, uint32_t b) {
    const float r = __builtin_amdgcn_sqrtf(-1.3862943611198906f *    // -2 ln 2
                                           __builtin_amdgcn_logf(u01f_open0(a)));
    return r * __builtin_amdgcn_cosf(u01f(b));
}
// Both Box-Muller normals of one word pair (independent N(0, 1) variates).
__device__ inline float2 normal2_f(uint32_t a, uint32_t b) {
    const float r = __builtin_amdgcn_sqrtf(-1.3862943611198906f *
                                           __builtin_amdgcn_logf(u01f_open0(a)));
    const float t = u01f(b);
    return make_float2(r * __builtin_amdgcn_cosf(t), r * __builtin_amdgcn_sinf(t));
}

// log(k!) in float32: the correctly rounded value for k < 10, Stirling's
// series with two correction terms for k >= 10 (truncation error < 5e-9).
__device__ inline float log_factorial(float kf) {
    const float m = kf + 1.0f, r = __builtin_amdgcn_rcpf(m);
    const float st = (m - 0.5f) * ln_hw(m) - m + 0x1.d67f1cp-1f +   // 0.5 log(2 pi)
                     r * (1.0f / 12.0f - (r * r) * (1.0f / 360.0f));
    const int k = (int)kf;
    float t = 0.0f;
    t = k == 2 ? 0x1.62e430p-1f : t;
    t = k == 3 ? 0x1.cab0c0p+0f : t;
    t = k == 4 ? 0x1.96ca78p+1f : t;
    t = k == 5 ? 0x1.326644p+2f : t;
    t = k == 6 ? 0x1.a51274p+2f : t;
    t = k == 7 ? 0x1.10ce20p+3f : t;
    t = k == 8 ? 0x1.5358e8p+3f : t;
    t = k == 9 ? 0x1.99a892p+3f : t;
    return kf >= 10.0f ? st : t;
}

// Poisson for lam >= 12: Hormann's PTRS (transformed rejection with squeeze,
// Insurance: Mathematics and Economics 12, 1993), two proposals per Philox
// block.  log(lam) stays the accurate logf (it is multiplied by k in the
// acceptance test); the other logarithms run on v_log_f32, square root and
// reciprocals on v_sqrt_f32 / v_rcp_f32 (~1 ulp: the constants of the method
// are themselves fitted to 3-4 digits).
__device__ inline int poisson_ptrs(const PixelRng& R, float lam) {
    const float slam = __builtin_amdgcn_sqrtf(lam), loglam = logf(lam);
    const float b = 0.931f + 2.53f * slam;
    const float aa = -0.059f + 0.02483f * b;
    const float log_invalpha = ln_hw(1.1239f + 1.1328f * __builtin_amdgcn_rcpf(b - 3.4f));
    const float vr = 0.9277f - 3.6224f * __builtin_amdgcn_rcpf(b - 2.0f);
    for (uint32_t blk = 1; blk <= 256; ++blk) {
        const smx::u32x4 r = R.block(blk);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const float U = u01f(h ? r.z : r.x) - 0.5f;
            const float V = u01f(h ? r.w : r.y);
            const float us = 0.5f - fabsf(U);
            const float kf = floorf((2.0f * aa * __builtin_amdgcn_rcpf(us) + b) * U + lam + 0.43f);
            if (us >= 0.07f && V <= vr) return (int)kf;
            if (kf < 0.0f || (us < 0.013f && V > us)) continue;
            if (ln_hw(V) + log_invalpha - ln_hw(aa * __builtin_amdgcn_rcpf(us * us) + b) <=
                -lam + kf * loglam - log_factorial(kf))
                return (int)kf;
        }
    }
    return (int)floorf(lam + 0.5f);
}

// Poisson(lam) by inversion: the smallest k with u < P(X <= k) (lam < 12).
__device__ inline int poisson_inversion(float lam, float u) {
    int k = 0;
    float p = __builtin_amdgcn_exp2f(-1.4426950408889634f * lam), cum = p;
    while (u > cum && k < 4096) {
        ++k;
        p *= lam * __builtin_amdgcn_rcpf((float)k);
        cum += p;
    }
    return k;
}

// Wave-aggregated append to an LDS queue: one LDS atomic per wave; returns
// the slot of an appending lane (-1 for the others).  Call with the wave's
// active lanes converged.
__device__ inline int queue_push(bool pred, int* counter) {
    const unsigned long long m = __ballot(pred);
    if (m == 0ull) return -1;
    const int leader = __ffsll((long long)m) - 1;
    const int lane = (int)__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, __popcll(m));
    base = __builtin_amdgcn_readlane(base, leader);   // leader is uniform
    const int rank = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                    __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
    return pred ? base + rank : -1;
}

// Marsaglia-Tsang iterations it0, it0 + 1, ... of Gamma(n, 1) on Philox
// blocks blk, blk + 1, ... (the loop of gamma_mt without the first-draw
// shortcut).
__device__ inline float gamma_mt_from(const PixelRng& R, int n, uint32_t& blk, int it0) {
    const float d = (float)n - 1.0f / 3.0f;
    const float c = __builtin_amdgcn_rsqf(9.0f * d);   // v_rsq_f32, ~1 ulp
    for (int it = it0; it < 256; ++it) {
        smx::u32x4 r = R.block(blk++);
        const float x = normal_f(r.x, r.y);
        const float u = u01f_open0(r.z);
        float v = 1.0f + c * x;
        if (v <= 0.0f) continue;
        v = v * v * v;
        if (u < 1.0f - 0.0331f * (x * x) * (x * x)) return d * v;
        if (ln_hw(u) < 0.5f * x * x + d * (1.0f - v + ln_hw(v))) return d * v;
    }
    return d;
}

// Work queues of one block, carved from a caller-provided LDS buffer of
// lds_bytes(npix) bytes (npix: pixels of the block, at most 32768).
struct BlockLds {
    int* k;        // Poisson count per pixel
    float* lam;    // expectation (PTRS and inversion items)
    float* x0;     // spare normal of block 0
    float* u0;     // spare uniform of block 0
    float* S;      // inversion uniform (phases 1-2), then the EM-register output
    short* q;      // PTRS queue, then the Gamma queue
    short* r;      // inversion queue (phases 1-2), then the Gamma retry queue
                   // (first draw rejected: flag 1 << 14)
    int* n;        // PTRS, Gamma, retry and inversion queue lengths
};

__host__ __device__ constexpr int lds_bytes(int npix) { return npix * 24 + 16; }

__device__ inline BlockLds carve(char* base, int npix) {
    BlockLds L;
    L.k = (int*)base;
    L.lam = (float*)(base + 4 * npix);
    L.x0 = (float*)(base + 8 * npix);
    L.u0 = (float*)(base + 12 * npix);
    L.S = (float*)(base + 16 * npix);
    L.q = (short*)(base + 20 * npix);
    L.r = (short*)(base + 22 * npix);
    L.n = (int*)(base + 24 * npix);
    return L;
}

// EMCCD register and readout noise for the NT * PPT pixels of one block
// (pixel loc = threadIdx.x + j * NT, j < PPT; in[j] false: no such pixel),
// per pixel: block 0 of its Philox stream gives the readout normal, a spare
// normal and a spare uniform for the first Gamma proposal, and the Poisson
// uniform; k = Poisson(E) by inversion below 12, else PTRS on blocks 1..;
// the EM register S = round(Gamma(k, emgain)) by Marsaglia-Tsang (every
// shape k >= 1 on one code path), redrawn below 0.5, on blocks 2^16..  Every
// draw comes from a fixed block of the pixel's own stream (pid(loc)), so the
// result does not depend on which thread or kernel computes it.
//
// Work queues: with one thread per pixel a wave would run the inversion
// loop, PTRS and every Marsaglia-Tsang retry whenever any of its 64 lanes
// needs them; here the pixels that take PTRS (lam >= 12) and those whose
// inversion is not decided by its first term are appended to LDS queues and
// drained by the whole block, then the pixels with k >= 1 take their first
// Gamma proposal from the block-0 spares, and a rejected proposal goes to a
// retry queue, so a wave runs a retry only for pixels that need one.
// Returns S (0 where k = 0) and the readout noise per pixel.  Contains
// __syncthreads(): every thread of the block calls it.
template <int NT, int PPT, typename PidFn>
__device__ __forceinline__ void emccd_block(const BlockLds& L, const float (&E_in)[PPT],
                                            const bool (&in)[PPT], PidFn pid, uint32_t frame,
                                            uint32_t k0, uint32_t k1,
                                            const smx_detector_params_t& P, float (&S_out)[PPT],
                                            float (&noise)[PPT]) {
    const int t = threadIdx.x;
    const float emgain = (float)P.emgain;
    if (t < 4) L.n[t] = 0;
    __syncthreads();

    // phase 1: block 0 of every pixel; k = 0 where the inversion's first term
    // decides, the other pixels below 12 queued for the inversion and those
    // from 12 up for PTRS (waves then drain pixels of similar work)
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        const int loc = t + j * NT;
        noise[j] = 0.0f;
        float E = 0.0f, x0 = 0.0f, u0 = 1.0f;
        bool ptrs = false, inv = false;
        if (in[j]) {
            E = E_in[j];
            const PixelRng R{pid(loc), frame, k0, k1};
            const smx::u32x4 r0 = R.block(0);
            const float2 nn = normal2_f(r0.x, r0.y);
            if (P.readout_noise > 0.0) noise[j] = (float)P.readout_noise * nn.x;
            x0 = nn.y;
            u0 = u01f_open0(r0.w);
            if (E > 0.0f) {
                if (E < 12.0f) {
                    const float u = u01f(r0.z);
                    // the inversion's first test (k = 0), as in poisson_inversion
                    inv = u > __builtin_amdgcn_exp2f(-1.4426950408889634f * E);
                    if (inv) L.S[loc] = u;
                } else {
                    ptrs = true;
                }
            }
        }
        L.k[loc] = 0;   // queued pixels get theirs in phase 2
        L.lam[loc] = E;
        L.x0[loc] = x0;
        L.u0[loc] = u0;
        const int slot = queue_push(ptrs, &L.n[0]);
        if (ptrs) L.q[slot] = (short)loc;
        const int islot = queue_push(inv, &L.n[3]);
        if (inv) L.r[islot] = (short)loc;
    }
    __syncthreads();

    // phase 2: PTRS (lam >= 12), then the inversions, one queued pixel per lane
    const int np = L.n[0], ni = L.n[3];
    for (int q = t; q < np + ni; q += NT) {
        if (q < np) {
            const int loc = L.q[q];
            const PixelRng R{pid(loc), frame, k0, k1};
            L.k[loc] = poisson_ptrs(R, L.lam[loc]);
        } else {
            const int loc = L.r[q - np];
            L.k[loc] = poisson_inversion(L.lam[loc], L.S[loc]);
        }
    }
    __syncthreads();

    // phase 3: Gamma queue (k >= 1)
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        const int loc = t + j * NT;
        const bool g = L.k[loc] > 0;
        const int slot = queue_push(g, &L.n[1]);
        if (g) L.q[slot] = (short)loc;
    }
    __syncthreads();

    // phase 4: first Marsaglia-Tsang proposal from the block-0 spares; a
    // rejected proposal or a result below 0.5 goes to the retry queue
    const int ng = L.n[1];
    for (int q0 = 0; q0 < ng; q0 += NT) {
        const int q = q0 + t;
        bool retry = false;
        int loc = 0;
        bool rejected = false;
        if (q < ng) {
            loc = L.q[q];
            const int k = L.k[loc];
            const float d = (float)k - 1.0f / 3.0f;
            const float c = __builtin_amdgcn_rsqf(9.0f * d);   // v_rsq_f32, ~1 ulp
            const float x = L.x0[loc], u = L.u0[loc];
            float v = 1.0f + c * x;
            bool acc = false;
            if (v > 0.0f) {
                v = v * v * v;
                acc = (u < 1.0f - 0.0331f * (x * x) * (x * x)) ||
                      (ln_hw(u) < 0.5f * x * x + d * (1.0f - v + ln_hw(v)));
            }
            const float sv = d * v * emgain;
            rejected = !acc;
            if (acc && sv >= 0.5f) {
                float S = floorf(sv + 0.5f);
                if (S < 1.0f) S = 1.0f;
                L.S[loc] = S;
            } else {
                retry = true;
            }
        }
        const int slot = queue_push(retry, &L.n[2]);
        if (retry) L.r[slot] = (short)(loc | (rejected ? (1 << 14) : 0));
    }
    __syncthreads();

    // phase 5: retries, continuing the pixel's stream at block 2^16
    const int nr = L.n[2];
    for (int q = t; q < nr; q += NT) {
        const int ent = L.r[q];
        const int loc = ent & ((1 << 14) - 1);
        const int k = L.k[loc];
        const PixelRng R{pid(loc), frame, k0, k1};
        uint32_t blk = 1u << 16;
        float s;
        if (ent & (1 << 14)) {
            // first draw rejected: its Gamma call goes on with iteration 1
            s = gamma_mt_from(R, k, blk, 1) * emgain;
        } else {
            const float d = (float)k - 1.0f / 3.0f;
            const float c = __builtin_amdgcn_rsqf(9.0f * d);   // v_rsq_f32, ~1 ulp
            const float v = 1.0f + c * L.x0[loc];
            s = d * (v * v * v) * emgain;   // accepted, below 0.5
        }
        for (int guard = 1; guard < 64 && !(s >= 0.5f); ++guard)
            s = gamma_mt_from(R, k, blk, 0) * emgain;
        float S = floorf(s + 0.5f);
        if (S < 1.0f) S = 1.0f;
        L.S[loc] = S;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        const int loc = t + j * NT;
        S_out[j] = L.k[loc] > 0 ? L.S[loc] : 0.0f;
    }
}

// ADC of one pixel in float32 (_epifm.py:1475-1484 order: fullwell clip,
// gain, offset, range clip).
__device__ inline float adc_f32(float S, float noise, const smx_detector_params_t& P, float gain,
                                float offset) {
    float x = S + noise;
    x = fminf(x, (float)P.fullwell);
    x = x * __builtin_amdgcn_rcpf(gain) + offset;
    return fminf(fmaxf(x, 0.0f), (float)P.adc_max);
}

// Expectation E = QE * (photons + background) in float32, as the Philox
// detector paths compute it.
__device__ inline float expectation_f32(double photons, const smx_detector_params_t& P) {
    float ph = (float)photons;
    if (P.background_switch) ph += (float)P.background;
    return (float)P.qe * ph;
}

}  // namespace emccd
}  // namespace smx

#endif  // SMX_EMCCD_BLOCK_HPP
