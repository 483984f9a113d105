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
//              sums the pmf by chunks, pass 2 finds the chunk and then the
//              term where the cumulative mass passes u * total.  Readout
//              noise comes in pre-drawn.
//  * philox  : statistically equivalent sampler driven by Philox4x32-10:
//              n ~ Poisson(E), S = round(Gamma(n, emgain)) (S >= 1 when n >= 1),
//              noise ~ N(0, readout_noise^2).  O(1) per pixel.
#include "common.hpp"
#include "emccd_block.hpp"
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

// Chunked form of the parity draw: the support's terms in chunks of PCL
// consecutive S, PJ consecutive terms per lane; pass 1 keeps every chunk's
// sum (lane sums, then one wave reduction per chunk) and the total is their
// sum in chunk order, so pass 2 finds the chunk where the cumulative mass
// passes u T from the chunk sums and evaluates only that chunk's terms (one
// support plus a chunk instead of about one and a half supports).  Supports
// of more than PCH chunks take the unchunked passes.
constexpr int PJ = 8;
constexpr int PCL = 64 * PJ;
constexpr int PCH = 192;

__device__ inline void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Pass 1, chunked: ch[c] = sum of terms PCL c .. PCL c + PCL - 1 of the
// support (c < nch); returns the total, the chunk sums added in order.
__device__ inline double emccd_chunk_sums(double E, double a, long long s_min, long long len,
                                          long long nch, int lane, double* ch) {
    double T = 0.0;
    for (long long c = 0; c < nch; ++c) {
        const long long k0 = PCL * c + PJ * lane;
        double p = 0.0;
        for (int j = 0; j < PJ; ++j)
            if (k0 + j < len) p += emccd_term(s_min + k0 + j, E, a);
        const double v = wave_sum(p);
        if (lane == 0) ch[c] = v;
        T += v;
    }
    wave_sync_lds();
    return T;
}

// Pass 2: the first support index whose cumulative mass exceeds thr (the
// last index when rounding leaves none): the chunk from the chunk sums, then
// the lane from a prefix scan of the lanes' sums of the chunk's terms, then
// the term among that lane's PJ (continuing into the next chunks if the
// recomputed terms have not yet passed thr at the chunk's end).
__device__ inline long long emccd_find(double E, double a, long long s_min, long long len,
                                       long long nch, const double* ch, double thr, int lane) {
    double carry = 0.0;
    long long c0 = nch;
    for (long long b = 0; b < nch; b += 64) {
        const long long c = b + lane;
        double v = c < nch ? ch[c] : 0.0;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const double o = __shfl_up(v, off, 64);
            if (lane >= off) v += o;
        }
        const double cum = carry + v;
        const double up = __shfl_up(cum, 1, 64);
        const double before = lane == 0 ? carry : up;   // mass before chunk c
        const unsigned long long hit = __ballot(c < nch && cum > thr);
        if (hit) {
            const int l = __ffsll((long long)hit) - 1;
            c0 = b + l;
            carry = __shfl(before, l, 64);
            break;
        }
        carry = __shfl(cum, 63, 64);
    }
    if (c0 >= nch) return len - 1;
    for (long long cb = PCL * c0; cb < len; cb += PCL) {
        const long long k0 = cb + PJ * lane;
        double p = 0.0;
        for (int j = 0; j < PJ; ++j)
            if (k0 + j < len) p += emccd_term(s_min + k0 + j, E, a);
        double v = p;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const double o = __shfl_up(v, off, 64);
            if (lane >= off) v += o;
        }
        const double cum = carry + v;
        const double up = __shfl_up(cum, 1, 64);
        const double before = lane == 0 ? carry : up;   // mass before this lane's terms
        const unsigned long long hit = __ballot(k0 < len && cum > thr);
        if (hit) {
            // lane l's PJ terms, one per lane j < PJ, scanned from the mass
            // before them; rounding may leave the scan at or below thr at the
            // end: then lane l's last term
            const int l = __ffsll((long long)hit) - 1;
            const long long kl = cb + PJ * l;
            const double bl = __shfl(before, l, 64);
            double t = lane < PJ && kl + lane < len ? emccd_term(s_min + kl + lane, E, a) : 0.0;
#pragma unroll
            for (int off = 1; off < PJ; off <<= 1) {
                const double o = __shfl_up(t, off, 64);
                if (lane >= off) t += o;
            }
            const unsigned long long in = __ballot(lane < PJ && bl + t > thr);
            const long long found = kl + (in ? (long long)(__ffsll((long long)in) - 1) : PJ - 1);
            return found < len - 1 ? found : len - 1;
        }
        carry = __shfl(cum, 63, 64);
    }
    return len - 1;
}

// The pmf of one expectation value shared by many pixels: T_out[0] = total,
// T_out[1 ..] = its chunk sums (when the support has at most PCH chunks).
__global__ void emccd_common_total_kernel(double E, int g, double* __restrict__ T_out) {
    int lane = threadIdx.x & 63;
    double a = 1.0 / (double)g;
    double T = 0.0;
    if (E > 0.0) {
        long long s_min, s_max;
        emccd_support(E, g, s_min, s_max);
        const long long len = s_max - s_min, nch = (len + PCL - 1) / PCL;
        if (nch <= PCH) T = emccd_chunk_sums(E, a, s_min, len, nch, lane, T_out + 1);
        else T = emccd_total(E, g, a, lane);
    }
    if (lane == 0) *T_out = T;
}

__global__ void __launch_bounds__(256) emccd_parity_kernel(
        int n, const double* __restrict__ E_arr, const double* __restrict__ noise,
        const double* __restrict__ uniforms, const int* __restrict__ uidx,
        const double* __restrict__ gain_map, const double* __restrict__ offset_map,
        smx_detector_params_t P, double E_common, const double* __restrict__ T_common,
        double* __restrict__ adc_out, double* __restrict__ signal_out, int* __restrict__ status) {
    __shared__ double s_ch[4][PCH];   // chunk sums, one pixel per wave
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
        const long long len = s_max - s_min, nch = (len + PCL - 1) / PCL;
        const bool common = T_common && E == E_common;
        const bool chunked = nch <= PCH;
        double* const ch = s_ch[threadIdx.x >> 6];
        double T;
        if (common) T = *T_common;
        else if (chunked) T = emccd_chunk_sums(E, a, s_min, len, nch, lane, ch);
        else T = emccd_total(E, g, a, lane);
        if (!(T > 0.0) || !(T < INFINITY)) {
            // the reference's pmf overflows here (numpy.random.choice raises ValueError)
            if (lane == 0) atomicOr(status, 4);
        } else if (chunked) {
            const double thr = uniforms[uidx[pix]] * T;
            S = (double)(s_min + emccd_find(E, a, s_min, len, nch, common ? T_common + 1 : ch,
                                            thr, lane));
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
// Philox samplers: emccd_block.hpp (shared with the fused splat + detector
// tile pass).  Float32 arithmetic: these paths are statistical, not
// bit-parity paths.
// ---------------------------------------------------------------------------
using namespace smx::emccd;

// EMCCD + ADC on Philox streams: a block of 256 threads owns 1024 consecutive
// pixels (4 per thread, coalesced) and runs emccd_block's work queues over
// them.
constexpr int EQ_THREADS = 256, EQ_PPT = 4, EQ_PIX = EQ_THREADS * EQ_PPT;

template <typename OutT>
__global__ void __launch_bounds__(EQ_THREADS) emccd_queue_kernel(
        int n, const double* __restrict__ photons, const double* __restrict__ gain_map,
        const double* __restrict__ offset_map, smx_detector_params_t P, uint64_t seed,
        uint32_t frame, uint32_t pix_base, const uint32_t* __restrict__ frame_dev,
        double* __restrict__ E_out, OutT* __restrict__ adc_out) {
    __shared__ __align__(16) char s_lds[lds_bytes(EQ_PIX)];
    const BlockLds L = carve(s_lds, EQ_PIX);
    const int t = threadIdx.x;
    const long long base = (long long)blockIdx.x * EQ_PIX;
    if (frame_dev) frame += *frame_dev;
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    float E[EQ_PPT];
    bool in[EQ_PPT];
#pragma unroll
    for (int j = 0; j < EQ_PPT; ++j) {
        const long long i = base + t + j * EQ_THREADS;
        in[j] = i < n;
        E[j] = 0.0f;
        if (in[j]) {
            E[j] = expectation_f32(photons[i], P);
            if (E_out) E_out[i] = (double)E[j];
        }
    }
    float S[EQ_PPT], noise[EQ_PPT];
    emccd_block<EQ_THREADS, EQ_PPT>(
            L, E, in, [&](int loc) { return pix_base + (uint32_t)(base + loc); }, frame, k0, k1,
            P, S, noise);
    // readout noise and ADC (_epifm.py:1475-1484)
#pragma unroll
    for (int j = 0; j < EQ_PPT; ++j) {
        const long long i = base + t + j * EQ_THREADS;
        if (!in[j]) continue;
        const float gain = (float)(gain_map ? gain_map[i] : P.gain);
        const float offset = (float)(offset_map ? offset_map[i] : P.offset);
        adc_out[i] = (OutT)adc_f32(S[j], noise[j], P, gain, offset);
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

size_t smx_emccd_common_elems(void) { return 1 + PCH; }

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
