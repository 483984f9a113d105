// Detector model: photons -> expectation E -> EMCCD signal + readout noise ->
// ADC counts.
//
// Reference: src/scopyon/_epifm.py:1430-1470 (__detector_output),
//            _epifm.py:355-416 (EMCCD), _epifm.py:1472-1484 (ADC).
//
// Two signal paths:
//  * parity  : the reference's discrete pmf on S in [s_min, s_max) drawn by
//              inverse CDF with a given uniform per pixel (the uniform numpy's
//              RandomState.choice would consume).  One wave per pixel: pass 1
//              sums the pmf, pass 2 scans it until the cumulative mass passes
//              u * total.  Readout noise comes in pre-drawn.
//  * philox  : statistically equivalent sampler driven by Philox4x32-10:
//              n ~ Poisson(E), S = round(Gamma(n, emgain)) (S >= 1 when n >= 1),
//              noise ~ N(0, readout_noise^2).  O(1) per pixel.
#include "common.hpp"
#include "scopyon_hip_types.h"

#include <hipcub/hipcub.hpp>

namespace {

__device__ inline double adc_counts(double x, const smx_detector_params_t& P, double gain,
                                    double offset) {
    // _epifm.py:1475-1484, same order, no rounding
    if (x > P.fullwell) x = P.fullwell;
    x = x / gain;
    x = x + offset;
    if (x > P.adc_max) x = P.adc_max;
    if (x < 0.0) x = 0.0;
    return x;
}

__device__ inline double expectation(double photons, const smx_detector_params_t& P) {
    if (P.background_switch) photons = photons + P.background;
    return P.qe * photons;
}

// Scaled compound Poisson-Gamma pmf term (_epifm.py:364-388).
__device__ inline double emccd_term(long long S, double E, double a) {
    if (S == 0) return 1.0 / (2.0 * a * E);
    double X = a * (double)S;
    double Y = 2.0 * sqrt(E * X);
    return 1.0 / Y * exp(-X + Y) * smx::i1e(Y);
}

__device__ inline void emccd_support(double E, int g, long long& s_min, long long& s_max) {
    double sigma = sqrt(E) * 5.0 + 10.0;
    long long lo = (long long)g * (long long)(E - sigma);   // int() truncates toward 0
    s_min = lo > 0 ? lo : 0;
    s_max = (long long)g * (long long)(E + sigma);
}

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Pass 1 of the parity draw: total mass of the pmf for E (wave-cooperative).
__device__ inline double emccd_total(double E, int g, double a, int lane) {
    long long s_min, s_max;
    emccd_support(E, g, s_min, s_max);
    double acc = 0.0;
    for (long long S = s_min + lane; S < s_max; S += 64) acc += emccd_term(S, E, a);
    return wave_sum(acc);
}

__global__ void prepare_kernel(int n, const double* __restrict__ photons, smx_detector_params_t P,
                               double* __restrict__ E_out, int* __restrict__ flag) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double E = expectation(photons[i], P);
    E_out[i] = E;
    if (flag) flag[i] = E > 0.0 ? 1 : 0;
}

// T_common = total pmf mass for one expectation value shared by many pixels.
__global__ void emccd_common_total_kernel(double E, int g, double* __restrict__ T_out) {
    int lane = threadIdx.x & 63;
    double a = 1.0 / (double)g;
    double T = E > 0.0 ? emccd_total(E, g, a, lane) : 0.0;
    if (lane == 0) *T_out = T;
}

__global__ void __launch_bounds__(256) emccd_parity_kernel(
        int n, const double* __restrict__ E_arr, const double* __restrict__ noise,
        const double* __restrict__ uniforms, const int* __restrict__ uidx,
        const double* __restrict__ gain_map, const double* __restrict__ offset_map,
        smx_detector_params_t P, double E_common, const double* __restrict__ T_common,
        double* __restrict__ adc_out, double* __restrict__ signal_out, int* __restrict__ status) {
    const int lane = threadIdx.x & 63;
    const int pix = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (pix >= n) return;
    const int g = P.emgain;
    const double a = 1.0 / (double)g;
    const double E = E_arr[pix];
    double S = 0.0;
    if (!(E <= 1.0e4)) {
        // NaN (int(nan) raises in the reference) or far beyond E ~ 709.8, where
        // exp(-X + Y) overflows and numpy.random.choice raises ValueError
        if (lane == 0) atomicOr(status, 4);
    } else if (E > 0.0) {
        long long s_min, s_max;
        emccd_support(E, g, s_min, s_max);
        const long long len = s_max - s_min;
        double T = (T_common && E == E_common) ? *T_common : emccd_total(E, g, a, lane);
        if (!(T > 0.0) || !(T < INFINITY)) {
            // the reference's pmf overflows here (numpy.random.choice raises ValueError)
            if (lane == 0) atomicOr(status, 4);
        } else {
            const double thr = uniforms[uidx[pix]] * T;
            double carry = 0.0;
            long long idx = len - 1;
            for (long long c0 = 0; c0 < len; c0 += 64) {
                long long k = c0 + lane;
                double v = k < len ? emccd_term(s_min + k, E, a) : 0.0;
#pragma unroll
                for (int off = 1; off < 64; off <<= 1) {
                    double o = __shfl_up(v, off, 64);
                    if (lane >= off) v += o;
                }
                double cum = carry + v;
                unsigned long long hit = __ballot(k < len && cum > thr);
                if (hit) {
                    long long found = c0 + (long long)(__ffsll((long long)hit) - 1);
                    idx = found < len - 1 ? found : len - 1;
                    break;
                }
                carry = __shfl(cum, 63, 64);
            }
            S = (double)(s_min + idx);
        }
    }
    if (lane == 0) {
        double x = S + (noise ? noise[pix] : 0.0);
        double gain = gain_map ? gain_map[pix] : P.gain;
        double offset = offset_map ? offset_map[pix] : P.offset;
        adc_out[pix] = adc_counts(x, P, gain, offset);
        if (signal_out) signal_out[pix] = S;
    }
}

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
        return smx::philox4x32_10(smx::u32x4{id, frame, 0x0E3CCD01u, b}, k0, k1);
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

// Gamma(n, 1), n >= 1: Marsaglia-Tsang (ACM TOMS 26(3), 2000), one Philox
// block per iteration (normal from two words, uniform from the third).
// With `first`, the first iteration takes its normal x0 and uniform u0 from
// the caller (the spare Box-Muller normal and spare word of the pixel's block
// 0), so an accepted first proposal -- the common case -- costs no further
// Philox block.
__device__ inline float gamma_mt(const PixelRng& R, int n, uint32_t& blk, bool first, float x0,
                                float u0) {
    const float d = (float)n - 1.0f / 3.0f;
    const float c = __builtin_amdgcn_rsqf(9.0f * d);   // v_rsq_f32, ~1 ulp
    for (int it = 0; it < 256; ++it) {
        float x, u;
        if (it == 0 && first) {
            x = x0;
            u = u0;
        } else {
            smx::u32x4 r = R.block(blk++);
            x = normal_f(r.x, r.y);
            u = u01f_open0(r.z);
        }
        float v = 1.0f + c * x;
        if (v <= 0.0f) continue;
        v = v * v * v;
        if (u < 1.0f - 0.0331f * (x * x) * (x * x)) return d * v;
        if (ln_hw(u) < 0.5f * x * x + d * (1.0f - v + ln_hw(v))) return d * v;
    }
    return d;
}

// EMCCD (_epifm.py:1448-1454, noise and register of _epifm.py:355-416) on
// Philox streams, float32.  Per pixel: block 0 gives the readout normal, a
// spare normal and a spare uniform for the first Gamma proposal, and the
// Poisson uniform; k = Poisson(E) by inversion below 12, else PTRS on blocks
// 1..; the EM register S = round(Gamma(k, emgain)) by Marsaglia-Tsang (every
// shape k >= 1 on one code path), redrawn below 0.5, on blocks 2^16..; then
// ADC.  Every draw comes from a fixed block of the pixel's own stream, so the
// result does not depend on which thread computes it.
//
// Work queues: a block of 256 threads owns 1024 pixels (4 per thread,
// coalesced).  With one thread per pixel a wave would run the inversion loop,
// PTRS and every Marsaglia-Tsang retry whenever any of its 64 lanes needs
// them; here the pixels that take PTRS (lam >= 12) and those with k >= 1 are
// appended to LDS queues and drained by the whole block, and a Gamma
// proposal rejected on the first (block-0 spare) draw goes to a retry queue,
// so a wave runs a retry only for pixels that need one.
constexpr int EQ_THREADS = 256, EQ_PPT = 4, EQ_PIX = EQ_THREADS * EQ_PPT;

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
    base = __shfl(base, leader);
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

template <typename OutT>
__global__ void __launch_bounds__(EQ_THREADS) emccd_queue_kernel(
        int n, const double* __restrict__ photons, const double* __restrict__ gain_map,
        const double* __restrict__ offset_map, smx_detector_params_t P, uint64_t seed,
        uint32_t frame, uint32_t pix_base, const uint32_t* __restrict__ frame_dev,
        double* __restrict__ E_out, OutT* __restrict__ adc_out) {
    __shared__ int s_k[EQ_PIX];          // Poisson count
    __shared__ float s_lam[EQ_PIX];      // expectation (PTRS items)
    __shared__ float s_x0[EQ_PIX];       // spare normal of block 0
    __shared__ float s_u0[EQ_PIX];       // spare uniform of block 0
    __shared__ float s_S[EQ_PIX];        // inversion uniform (phases 1-2), EM-register output
    __shared__ short s_q[EQ_PIX];        // PTRS queue, then the Gamma queue
    __shared__ short s_r[EQ_PIX];        // inversion queue (phases 1-2), then the Gamma
                                         // retry queue (first draw rejected: flag 1 << 14)
    __shared__ int s_n[4];               // PTRS, Gamma, retry and inversion queue lengths

    const int t = threadIdx.x;
    const long long base = (long long)blockIdx.x * EQ_PIX;
    if (t < 4) s_n[t] = 0;
    if (frame_dev) frame += *frame_dev;
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    const float emgain = (float)P.emgain;
    __syncthreads();

    // phase 1: block 0 of every pixel (readout normal, spare normal, spare
    // uniform, Poisson uniform); k = 0 where the inversion's first term
    // decides, the other pixels below 12 queued for the inversion and those
    // from 12 up for PTRS (waves then drain pixels of similar work)
    float noise[EQ_PPT];
#pragma unroll
    for (int j = 0; j < EQ_PPT; ++j) {
        const int loc = t + j * EQ_THREADS;
        const long long i = base + loc;
        const bool in = i < n;
        noise[j] = 0.0f;
        float E = 0.0f, x0 = 0.0f, u0 = 1.0f;
        bool ptrs = false, inv = false;
        if (in) {
            float ph = (float)photons[i];
            if (P.background_switch) ph += (float)P.background;
            E = (float)P.qe * ph;
            const PixelRng R{pix_base + (uint32_t)i, frame, k0, k1};
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
                    if (inv) s_S[loc] = u;
                } else {
                    ptrs = true;
                }
            }
            if (E_out) E_out[i] = (double)E;
        }
        s_k[loc] = 0;   // queued pixels get theirs in phase 2
        s_lam[loc] = E;
        s_x0[loc] = x0;
        s_u0[loc] = u0;
        const int slot = queue_push(ptrs, &s_n[0]);
        if (ptrs) s_q[slot] = (short)loc;
        const int islot = queue_push(inv, &s_n[3]);
        if (inv) s_r[islot] = (short)loc;
    }
    __syncthreads();

    // phase 2: PTRS (lam >= 12), then the inversions, one queued pixel per lane
    const int np = s_n[0], ni = s_n[3];
    for (int q = t; q < np + ni; q += EQ_THREADS) {
        if (q < np) {
            const int loc = s_q[q];
            const PixelRng R{pix_base + (uint32_t)(base + loc), frame, k0, k1};
            s_k[loc] = poisson_ptrs(R, s_lam[loc]);
        } else {
            const int loc = s_r[q - np];
            s_k[loc] = poisson_inversion(s_lam[loc], s_S[loc]);
        }
    }
    __syncthreads();

    // phase 3: Gamma queue (k >= 1)
#pragma unroll
    for (int j = 0; j < EQ_PPT; ++j) {
        const int loc = t + j * EQ_THREADS;
        const bool g = s_k[loc] > 0;
        const int slot = queue_push(g, &s_n[1]);
        if (g) s_q[slot] = (short)loc;
    }
    __syncthreads();

    // phase 4: first Marsaglia-Tsang proposal from the block-0 spares; a
    // rejected proposal or a result below 0.5 goes to the retry queue
    const int ng = s_n[1];
    for (int q0 = 0; q0 < ng; q0 += EQ_THREADS) {
        const int q = q0 + t;
        bool retry = false;
        int loc = 0;
        bool rejected = false;
        if (q < ng) {
            loc = s_q[q];
            const int k = s_k[loc];
            const float d = (float)k - 1.0f / 3.0f;
            const float c = __builtin_amdgcn_rsqf(9.0f * d);   // v_rsq_f32, ~1 ulp
            const float x = s_x0[loc], u = s_u0[loc];
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
                s_S[loc] = S;
            } else {
                retry = true;
            }
        }
        const int slot = queue_push(retry, &s_n[2]);
        if (retry) s_r[slot] = (short)(loc | (rejected ? (1 << 14) : 0));
    }
    __syncthreads();

    // phase 5: retries, continuing the pixel's stream at block 2^16
    const int nr = s_n[2];
    for (int q = t; q < nr; q += EQ_THREADS) {
        const int ent = s_r[q];
        const int loc = ent & ((1 << 14) - 1);
        const int k = s_k[loc];
        const PixelRng R{pix_base + (uint32_t)(base + loc), frame, k0, k1};
        uint32_t blk = 1u << 16;
        float s;
        if (ent & (1 << 14)) {
            // first draw rejected: its Gamma call goes on with iteration 1
            s = gamma_mt_from(R, k, blk, 1) * emgain;
        } else {
            const float d = (float)k - 1.0f / 3.0f;
            const float c = __builtin_amdgcn_rsqf(9.0f * d);   // v_rsq_f32, ~1 ulp
            const float v = 1.0f + c * s_x0[loc];
            s = d * (v * v * v) * emgain;   // accepted, below 0.5
        }
        for (int guard = 1; guard < 64 && !(s >= 0.5f); ++guard)
            s = gamma_mt_from(R, k, blk, 0) * emgain;
        float S = floorf(s + 0.5f);
        if (S < 1.0f) S = 1.0f;
        s_S[loc] = S;
    }
    __syncthreads();

    // phase 6: readout noise and ADC (_epifm.py:1475-1484)
#pragma unroll
    for (int j = 0; j < EQ_PPT; ++j) {
        const int loc = t + j * EQ_THREADS;
        const long long i = base + loc;
        if (i >= n) continue;
        const float S = s_k[loc] > 0 ? s_S[loc] : 0.0f;
        const float gain = (float)(gain_map ? gain_map[i] : P.gain);
        const float offset = (float)(offset_map ? offset_map[i] : P.offset);
        float x = S + noise[j];
        x = fminf(x, (float)P.fullwell);
        x = x * __builtin_amdgcn_rcpf(gain) + offset;
        x = fminf(fmaxf(x, 0.0f), (float)P.adc_max);
        adc_out[i] = (OutT)x;
    }
}

// searchsorted(cdf, u, side='right') clamped to the last index: the draw of
// RandomState.choice(values, p=p) given its uniform u (cdf as numpy builds it).
__device__ inline int table_index(const double* __restrict__ cdf, int m, double u) {
    int lo = 0, hi = m;   // first index with cdf[idx] > u
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (cdf[mid] <= u) lo = mid + 1;
        else hi = mid;
    }
    return lo < m ? lo : m - 1;
}

// CCD / CMOS with caller-drawn numbers (_epifm.py:1443-1469).
__global__ void adc_parity_kernel(int n, const double* __restrict__ signal,
                                  const double* __restrict__ noise_in,
                                  const double* __restrict__ tvals,
                                  const double* __restrict__ tcdf, int tn,
                                  const double* __restrict__ gain_map,
                                  const double* __restrict__ offset_map, smx_detector_params_t P,
                                  double* __restrict__ adc_out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double noise = 0.0;
    if (noise_in) noise = tn > 0 ? tvals[table_index(tcdf, tn, noise_in[i])] : noise_in[i];
    const double gain = gain_map ? gain_map[i] : P.gain;
    const double offset = offset_map ? offset_map[i] : P.offset;
    adc_out[i] = adc_counts(signal[i] + noise, P, gain, offset);
}

// CCD (Gaussian readout noise) or CMOS (readout noise from a value table)
// with Philox randomness: Poisson(E) photo-electrons + noise, float32 ADC.
template <typename OutT>
__global__ void photon_philox_kernel(int n, const double* __restrict__ photons,
                                     const double* __restrict__ gain_map,
                                     const double* __restrict__ offset_map,
                                     smx_detector_params_t P, const double* __restrict__ tvals,
                                     const double* __restrict__ tcdf, int tn, uint64_t seed,
                                     uint32_t frame, uint32_t pix_base,
                                     const uint32_t* __restrict__ frame_dev,
                                     double* __restrict__ E_out, OutT* __restrict__ adc_out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (frame_dev) frame += *frame_dev;
    float ph = (float)photons[i];
    if (P.background_switch) ph += (float)P.background;
    const float E = (float)P.qe * ph;
    PixelRng R{pix_base + (uint32_t)i, frame, (uint32_t)seed, (uint32_t)(seed >> 32)};
    const smx::u32x4 r0 = R.block(0);
    float noise = 0.0f;
    if (tn > 0) noise = (float)tvals[table_index(tcdf, tn, smx::u01_53(r0.x, r0.y))];
    else if (P.readout_noise > 0.0) noise = (float)P.readout_noise * normal_f(r0.x, r0.y);
    int k = 0;
    if (E > 0.0f) k = E < 12.0f ? poisson_inversion(E, u01f(r0.z)) : poisson_ptrs(R, E);
    const float gain = (float)(gain_map ? gain_map[i] : P.gain);
    const float offset = (float)(offset_map ? offset_map[i] : P.offset);
    float x = (float)k + noise;
    x = fminf(x, (float)P.fullwell);
    x = x * __builtin_amdgcn_rcpf(gain) + offset;
    x = fminf(fmaxf(x, 0.0f), (float)P.adc_max);
    if (E_out) E_out[i] = (double)E;
    adc_out[i] = (OutT)x;
}

}  // namespace

extern "C" {

int smx_detector_expectation(int n, const double* photons, const smx_detector_params_t* params,
                             double* E_out, int* flag_out, hipStream_t stream) {
    if (n < 0 || !params || (n > 0 && (!photons || !E_out))) {
        smx::set_error("smx_detector_expectation: invalid arguments (n=%d)", n);
        return SMX_EVALUE;
    }
    if (n == 0) return SMX_OK;
    hipLaunchKernelGGL(prepare_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, n, photons,
                       *params, E_out, flag_out);
    SMX_CHECK_LAUNCH("prepare_kernel");
    return SMX_OK;
}

size_t smx_scan_workspace_bytes(int n) {
    size_t bytes = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (int*)nullptr, (int*)nullptr, n);
    return bytes;
}

int smx_exclusive_scan_i32(int n, const int* in, int* out, void* workspace, size_t bytes,
                           hipStream_t stream) {
    if (n < 0 || (n > 0 && (!in || !out || !workspace))) {
        smx::set_error("smx_exclusive_scan_i32: invalid arguments (n=%d)", n);
        return SMX_EVALUE;
    }
    if (n == 0) return SMX_OK;
    size_t need = smx_scan_workspace_bytes(n);
    if (bytes < need) {
        smx::set_error("smx_exclusive_scan_i32: workspace too small (%zu < %zu)", bytes, need);
        return SMX_EVALUE;
    }
    if (smx::check_hip(hipcub::DeviceScan::ExclusiveSum(workspace, need, in, out, n, stream),
                       "hipcub ExclusiveSum"))
        return SMX_ERUNTIME;
    return SMX_OK;
}

int smx_emccd_common_total(double E, int emgain, double* T_out, hipStream_t stream) {
    if (emgain <= 0 || !T_out) {
        smx::set_error("smx_emccd_common_total: invalid arguments (emgain=%d)", emgain);
        return SMX_EVALUE;
    }
    hipLaunchKernelGGL(emccd_common_total_kernel, dim3(1), dim3(64), 0, stream, E, emgain, T_out);
    SMX_CHECK_LAUNCH("emccd_common_total_kernel");
    return SMX_OK;
}

int smx_emccd_parity(int n, const double* E, const double* noise, const double* uniforms,
                     const int* uidx, const double* gain_map, const double* offset_map,
                     const smx_detector_params_t* params, double E_common,
                     const double* T_common, double* adc_out, double* signal_out,
                     int* status_dev, hipStream_t stream) {
    if (n < 0 || !params || params->emgain <= 0 ||
        (n > 0 && (!E || !uniforms || !uidx || !adc_out || !status_dev))) {
        smx::set_error("smx_emccd_parity: invalid arguments (n=%d)", n);
        return SMX_EVALUE;
    }
    if (n == 0) return SMX_OK;
    hipLaunchKernelGGL(emccd_parity_kernel, dim3((n + 3) / 4), dim3(256), 0, stream, n, E, noise,
                       uniforms, uidx, gain_map, offset_map, *params, E_common, T_common, adc_out,
                       signal_out, status_dev);
    SMX_CHECK_LAUNCH("emccd_parity_kernel");
    return SMX_OK;
}

int smx_emccd_philox(int n, const double* photons, const double* gain_map,
                     const double* offset_map, const smx_detector_params_t* params,
                     uint64_t seed, uint32_t frame, uint32_t pix_base,
                     const uint32_t* frame_dev, double* E_out, void* adc_out,
                     int adc_is_float32, hipStream_t stream) {
    if (n < 0 || !params || params->emgain <= 0 || (n > 0 && (!photons || !adc_out))) {
        smx::set_error("smx_emccd_philox: invalid arguments (n=%d)", n);
        return SMX_EVALUE;
    }
    if (n == 0) return SMX_OK;
    const dim3 grid((unsigned)((n + EQ_PIX - 1) / EQ_PIX));
    if (adc_is_float32)
        hipLaunchKernelGGL(emccd_queue_kernel<float>, grid, dim3(EQ_THREADS), 0, stream, n,
                           photons, gain_map, offset_map, *params, seed, frame, pix_base,
                           frame_dev, E_out, (float*)adc_out);
    else
        hipLaunchKernelGGL(emccd_queue_kernel<double>, grid, dim3(EQ_THREADS), 0, stream, n,
                           photons, gain_map, offset_map, *params, seed, frame, pix_base,
                           frame_dev, E_out, (double*)adc_out);
    SMX_CHECK_LAUNCH("emccd_queue_kernel");
    return SMX_OK;
}

int smx_adc_parity(int n, const double* signal, const double* noise_in, const double* table_values,
                   const double* table_cdf, int table_n, const double* gain_map,
                   const double* offset_map, const smx_detector_params_t* params, double* adc_out,
                   hipStream_t stream) {
    if (n < 0 || !params || table_n < 0 || (table_n > 0 && (!table_values || !table_cdf)) ||
        (n > 0 && (!signal || !adc_out))) {
        smx::set_error("smx_adc_parity: invalid arguments (n=%d, table_n=%d)", n, table_n);
        return SMX_EVALUE;
    }
    if (n == 0) return SMX_OK;
    hipLaunchKernelGGL(adc_parity_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, n, signal,
                       noise_in, table_values, table_cdf, table_n, gain_map, offset_map, *params,
                       adc_out);
    SMX_CHECK_LAUNCH("adc_parity_kernel");
    return SMX_OK;
}

int smx_photon_detector_philox(int n, const double* photons, const double* gain_map,
                               const double* offset_map, const smx_detector_params_t* params,
                               const double* table_values, const double* table_cdf, int table_n,
                               uint64_t seed, uint32_t frame, uint32_t pix_base,
                               const uint32_t* frame_dev, double* E_out, void* adc_out,
                               int adc_is_float32, hipStream_t stream) {
    if (n < 0 || !params || table_n < 0 || (table_n > 0 && (!table_values || !table_cdf)) ||
        (n > 0 && (!photons || !adc_out))) {
        smx::set_error("smx_photon_detector_philox: invalid arguments (n=%d)", n);
        return SMX_EVALUE;
    }
    if (n == 0) return SMX_OK;
    if (adc_is_float32)
        hipLaunchKernelGGL(photon_philox_kernel<float>, dim3((n + 255) / 256), dim3(256), 0,
                           stream, n, photons, gain_map, offset_map, *params, table_values,
                           table_cdf, table_n, seed, frame, pix_base, frame_dev, E_out,
                           (float*)adc_out);
    else
        hipLaunchKernelGGL(photon_philox_kernel<double>, dim3((n + 255) / 256), dim3(256), 0,
                           stream, n, photons, gain_map, offset_map, *params, table_values,
                           table_cdf, table_n, seed, frame, pix_base, frame_dev, E_out,
                           (double*)adc_out);
    SMX_CHECK_LAUNCH("photon_philox_kernel");
    return SMX_OK;
}

}  // extern "C"
