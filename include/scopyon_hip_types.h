/* Plain-old-data structs shared by the scopyon HIP C-ABI and its kernels.
 * All lengths are SI metres, times seconds; all doubles are float64. */
#ifndef SCOPYON_HIP_TYPES_H
#define SCOPYON_HIP_TYPES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One field of view (camera frame) in a batch.
 * Image layout: pixel (i, j) of FOV f is out[pix_off + i * nh + j], i.e. the
 * reference's camera[i, j] with i along image_size[0] (u) and j along
 * image_size[1] (v) (src/scopyon/_epifm.py:1175-1177). */
typedef struct smx_fov_t {
    int32_t nw, nh;          /* image_size = (nw, nh) */
    int32_t tiles_r, tiles_c;/* ceil(nw / t), ceil(nh / t): tile side t = 32
                                (smx_splat_tile_size()), or the side given to
                                smx_splat_*_ts */
    int64_t pix_off;         /* first pixel of this frame in the batch buffer */
    int32_t tile_off;        /* first global tile id of this frame */
    int32_t reserved;
    double pixel_length;     /* detector pixel / magnification (_epifm.py:1176) */
    double half_w, half_h;   /* (nw * pixel_length) * 0.5, (nh * pixel_length) * 0.5 */
    double p0u, p0v;         /* focal point lateral coordinates (_epifm.py:1168) */
} smx_fov_t;

/* Per-emitter weight constants, all computed on the host in float64 with the
 * reference's operation order (src/scopyon/_epifm.py:1343-1428). */
typedef struct smx_weight_params_t {
    double amplitude;              /* snells_law() amplitude [photons/s/m^2] */
    double penetration_depth;      /* snells_law() depth [m] (inf for EPI) */
    double x_sec;                  /* ln(10) * abs_coeff * 0.1 / N_A */
    double quantum_yield;
    double one_minus_10_pow_minus_a; /* 1 - 10**(-A) */
    double spectral_psf_norm;      /* numpy.sum(fluoem_norm) * psf_normalization */
    double four_pi;                /* 4.0 * numpy.pi */
    double focal_depth;            /* p_0[0] */
    double depth_cutoff;           /* fluorophore.depth_cutoff */
    double res;                    /* 1e-9, depth key resolution */
    int32_t cutoff_slot;           /* table slot used for |depth| >= cutoff + res */
    int32_t reserved;
} smx_weight_params_t;

/* Detector constants (src/scopyon/_epifm.py:1430-1484). */
typedef struct smx_detector_params_t {
    double qe;                 /* detector.QE */
    double background;         /* effects.background.mean, 0 if switched off */
    int32_t background_switch;
    int32_t emgain;            /* detector.emgain (EMCCD) */
    double readout_noise;      /* detector.readout_noise (e-) */
    double fullwell;           /* ADC fullwell */
    double adc_max;            /* 2**bit - 1 */
    double offset;             /* scalar ADC offset (used when offset_map is NULL) */
    double gain;               /* scalar ADC gain (used when gain_map is NULL) */
} smx_detector_params_t;

/* Separable Gaussian image model in pixel units (the differentiable and the
 * matrix-core forms of the splat).  Emitter e at (x, y) (pixel i covers
 * [i, i + 1)), depth z and intensity I adds to pixel (i, j) of its frame
 *   I * exp(-|z| * inv_depth) * G(i; x) * G(j; y),
 *   G(i; x) = (erf((i + 1 - x) / (sqrt(2) sigma)) - erf((i - x) / (sqrt(2) sigma))) / 2,
 * over the window i in [floor(x) - radius, floor(x) + radius] (the Gaussian
 * radial profile of _epifm.py:128-134 integrated over each pixel; the depth
 * factor is the TIRF amplitude exp(-|depth| / penetration_depth) of
 * _epifm.py:1289).  The expectation is background + the sum over emitters. */
typedef struct smx_gauss_params_t {
    float sigma;        /* PSF width [px] */
    float inv_depth;    /* 1 / penetration depth [1/um]; 0 disables the depth factor */
    float background;   /* expected photons per pixel added to every frame */
    int32_t radius;     /* window half-width R (1 <= R <= 7): 2R + 1 pixels per axis */
} smx_gauss_params_t;

/* Radial x depth lookup-table PSF ("LUT mode", a deliberate deviation from
 * the reference's 1 nm depth keys and 1 nm cell sums, for 3D work at any
 * pixel size).  Per depth slice, K[qu][qv][du][dv] (float32) holds the
 * integral of the slice's PSF over footprint pixel (du, dv) when the
 * emitter sits at sub-pixel phase (qu / q, qv / q) of its pixel: pixel du
 * spans [du - c - qu / q, du - c + 1 - qu / q) pixels from the emitter.  The
 * PSF is the reference's cartesian kernel: the radial profile P(min(r,
 * r_max)) on the square |x|, |y| <= r_max, zero outside (_epifm.py:98-126),
 * with P from an FP64 Born-Wolf table on nb radial bins (_epifm.py:180-213).
 * Emitters take the bilinear combination of the four phases around their
 * exact position. */
typedef struct smx_lut_t {
    int32_t q;             /* phase steps per pixel (the table holds q + 1 per axis) */
    int32_t m;             /* footprint pixels per axis, 2 c + 1 */
    int32_t c;             /* footprint pixel holding the emitter (floor(r_max / pl) + 1) */
    int32_t nslices;       /* depth slices */
    int32_t ns;            /* build: midpoint cells per phase step per axis */
    int32_t reserved;
    double pixel_length;   /* detector pixel / magnification [m] */
    double r_max;          /* radial extent of the profile [m] */
} smx_lut_t;

#ifdef __cplusplus
}
#endif

#endif /* SCOPYON_HIP_TYPES_H */
