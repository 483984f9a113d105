// PSF tables: radial profiles (Born-Wolf / Gaussian, float64) and the
// summed-area table (SAT) of the (2n-1)^2 cartesian table for each profile.
//
// Reference: src/scopyon/_epifm.py:90-134 (radial profile, cartesian table),
//            _epifm.py:180-213 (Born-Wolf 100-point Riemann sum).
// The cartesian table is never stored: each SAT row pass evaluates its cells
// from the radial profile on the fly (same interpolation as interp1d).
#include "common.hpp"

namespace {

constexpr int kRho = 100;   // _epifm.py:191

// J0(x * alpha * rho_n) for every (radial sample, rho) pair; independent of z.
__global__ void bw_j0_kernel(int nr, double res, double alpha, double* __restrict__ j0tab) {
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nr * kRho) return;
    int ri = idx / kRho, n = idx % kRho;
    double r = (double)ri * res;                   // numpy.arange(0, cut, res)[ri]
    double rho = (double)(n + 1) * (1.0 / kRho);   // numpy.arange(1, 101) * drho
    j0tab[idx] = j0((r * alpha) * rho);
}

// psf[z][r] = |sum_n J0[r][n] * exp(-2i z gamma rho^2) rho drho|^2 * alpha^2/pi
__global__ void bw_profile_kernel(int nr, int nz, const double* __restrict__ depths,
                                  double gamma, double scale,
                                  const double* __restrict__ j0tab, double* __restrict__ out) {
    int ri = blockIdx.x * blockDim.x + threadIdx.x;
    int zi = blockIdx.y;
    if (ri >= nr || zi >= nz) return;
    double z = depths[zi];
    const double drho = 1.0 / kRho;
    double re = 0.0, im = 0.0;
    const double* row = j0tab + (size_t)ri * kRho;
    for (int n = 0; n < kRho; ++n) {
        double rho = (double)(n + 1) * drho;
        // -2 * 1j * z * gamma * rho * rho: imaginary part (((-2z) gamma) rho) rho
        double theta = (((-2.0 * z) * gamma) * rho) * rho;
        double s, c;
        sincos(theta, &s, &c);
        double yre = (c * rho) * drho, yim = (s * rho) * drho;
        re += yre * row[n];
        im += yim * row[n];
    }
    double mag = hypot(re, im);
    out[(size_t)zi * nr + ri] = (mag * mag) * scale;
}

__global__ void gaussian_profile_kernel(int nr, double res, double width, double denom,
                                        double* __restrict__ out) {
    int ri = blockIdx.x * blockDim.x + threadIdx.x;
    if (ri >= nr) return;
    double r = (double)ri * res;
    double q = r / width;
    out[ri] = exp(-0.5 * (q * q)) / denom;
}

// Linear interpolation of the radial profile at R (scipy interp1d 'linear':
// lo = clip(searchsorted_left(x, R), 1, n-1) - 1).
__device__ inline double interp_radial(const double* __restrict__ prof, int nr, double res,
                                       double R) {
    int m = (int)(R / res);
    if (m > nr) m = nr;
    if (m < 0) m = 0;
    // m := number of grid points x_k = k*res with x_k < R
    while (m < nr && (double)m * res < R) ++m;
    while (m > 0 && (double)(m - 1) * res >= R) --m;
    int k = m < 1 ? 1 : (m > nr - 1 ? nr - 1 : m);
    double xlo = (double)(k - 1) * res, xhi = (double)k * res;
    double ylo = prof[k - 1], yhi = prof[k];
    double slope = (yhi - ylo) / (xhi - xlo);
    return slope * (R - xlo) + ylo;
}

// One wave per table row: inclusive scan of the row's cells (evaluated from
// the radial profile) -> rowsum[row][col+1]; rowsum[row][0] = 0.
__global__ void __launch_bounds__(256) sat_row_kernel(int nr, double res, const double* __restrict__ prof,
                                                      int ntab, double* __restrict__ rowsum) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int row = blockIdx.x * 4 + wave;
    const int zi = blockIdx.y;
    if (row >= ntab) return;
    const int side = ntab + 1;
    const double* p = prof + (size_t)zi * nr;
    double* out = rowsum + ((size_t)zi * ntab + row) * side;
    const double rmax = (double)(nr - 1) * res;
    const double Y = (double)(row - (nr - 1)) * res;
    double carry = 0.0;
    if (lane == 0) out[0] = 0.0;
    for (int c0 = 0; c0 < ntab; c0 += 64) {
        int col = c0 + lane;
        double v = 0.0;
        if (col < ntab) {
            double X = (double)(col - (nr - 1)) * res;
            double R = sqrt(X * X + Y * Y);
            if (R > rmax) R = rmax;
            v = interp_radial(p, nr, res, R);
        }
        // wave-wide inclusive scan
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            double o = __shfl_up(v, off, 64);
            if (lane >= off) v += o;
        }
        double s = carry + v;
        if (col < ntab) out[col + 1] = s;
        carry = __shfl(s, 63, 64);
    }
}

// 64 columns per wave: running sum down the rows -> SAT in the chosen layout.
__global__ void __launch_bounds__(64) sat_col_kernel(int ntab, const double* __restrict__ rowsum,
                                                     smx::SatLayout L, size_t sat_stride,
                                                     double* __restrict__ sat) {
    const int side = ntab + 1;
    const int col = blockIdx.x * 64 + threadIdx.x;
    const int zi = blockIdx.y;
    if (col >= side) return;
    const double* rs = rowsum + (size_t)zi * ntab * side;
    double* out = sat + (size_t)zi * sat_stride;
    double acc = 0.0;
    out[L.index(0, col)] = 0.0;
    int row = 0;
    for (; row + 4 <= ntab; row += 4) {
        double a0 = rs[(size_t)(row + 0) * side + col];
        double a1 = rs[(size_t)(row + 1) * side + col];
        double a2 = rs[(size_t)(row + 2) * side + col];
        double a3 = rs[(size_t)(row + 3) * side + col];
        acc += a0; out[L.index(row + 1, col)] = acc;
        acc += a1; out[L.index(row + 2, col)] = acc;
        acc += a2; out[L.index(row + 3, col)] = acc;
        acc += a3; out[L.index(row + 4, col)] = acc;
    }
    for (; row < ntab; ++row) {
        acc += rs[(size_t)row * side + col];
        out[L.index(row + 1, col)] = acc;
    }
}

}  // namespace

extern "C" {

int smx_psf_radial_born_wolf(const double* depths, int nz, int nr, double res, double alpha,
                             double gamma, double scale, double* j0_work, double* out,
                             hipStream_t stream) {
    if (nz < 0 || nr < 2 || !depths || !j0_work || !out) {
        smx::set_error("smx_psf_radial_born_wolf: invalid arguments (nz=%d, nr=%d)", nz, nr);
        return SMX_EVALUE;
    }
    if (nz == 0) return SMX_OK;
    int total = nr * kRho;
    hipLaunchKernelGGL(bw_j0_kernel, dim3((total + 255) / 256), dim3(256), 0, stream, nr, res,
                       alpha, j0_work);
    SMX_CHECK_LAUNCH("bw_j0_kernel");
    hipLaunchKernelGGL(bw_profile_kernel, dim3((nr + 127) / 128, nz), dim3(128), 0, stream, nr,
                       nz, depths, gamma, scale, j0_work, out);
    SMX_CHECK_LAUNCH("bw_profile_kernel");
    return SMX_OK;
}

int smx_psf_radial_gaussian(int nr, double res, double width, double* out, hipStream_t stream) {
    if (nr < 2 || !(width > 0.0) || !out) {
        smx::set_error("smx_psf_radial_gaussian: invalid arguments (nr=%d, width=%g)", nr, width);
        return SMX_EVALUE;
    }
    double denom = 2.0 * 3.141592653589793 * width * width;
    hipLaunchKernelGGL(gaussian_profile_kernel, dim3((nr + 255) / 256), dim3(256), 0, stream, nr,
                       res, width, denom, out);
    SMX_CHECK_LAUNCH("gaussian_profile_kernel");
    return SMX_OK;
}

size_t smx_psf_sat_elems(int nr, int period) {
    if (nr < 2) return 0;
    return smx::make_sat_layout(2 * nr, period).elems();
}

size_t smx_psf_sat_workspace_bytes(int nr, int nz) {
    if (nr < 2 || nz < 0) return 0;
    size_t ntab = 2 * (size_t)nr - 1;
    return ntab * (ntab + 1) * (size_t)nz * sizeof(double);
}

int smx_psf_sat_build(const double* radial, int nz, int nr, double res, int period,
                      double* workspace, double* sat, hipStream_t stream) {
    if (nz < 0 || nr < 2 || period < 1 || !radial || !workspace || !sat) {
        smx::set_error("smx_psf_sat_build: invalid arguments (nz=%d, nr=%d, period=%d)", nz, nr,
                       period);
        return SMX_EVALUE;
    }
    if (nz == 0) return SMX_OK;
    int ntab = 2 * nr - 1;
    smx::SatLayout L = smx::make_sat_layout(ntab + 1, period);
    hipLaunchKernelGGL(sat_row_kernel, dim3((ntab + 3) / 4, nz), dim3(256), 0, stream, nr, res,
                       radial, ntab, workspace);
    SMX_CHECK_LAUNCH("sat_row_kernel");
    hipLaunchKernelGGL(sat_col_kernel, dim3((ntab + 1 + 63) / 64, nz), dim3(64), 0, stream, ntab,
                       workspace, L, L.elems(), sat);
    SMX_CHECK_LAUNCH("sat_col_kernel");
    return SMX_OK;
}

}  // extern "C"
