/* scopyon HIP C-ABI (libscopyon_hip.so, gfx950).
 *
 * The reference (ecell/scopyon, pure Python) has no FFI: its hot path sits
 * behind Python functions.  Each entry point below replaces one reference
 * loop or function; the citation names it (src/scopyon/<file>:<lines>).  The
 * Python package `scopyon_amd` binds these through ctypes and keeps the
 * reference's Python signatures on top.
 *
 * Conventions
 *   - All pointers are DEVICE pointers unless the comment says "host".
 *   - Every call is asynchronous on `stream`; nothing allocates device memory:
 *     scratch space comes in through `workspace` arguments sized by the
 *     matching *_workspace_bytes() query.
 *   - Return value: 0 on success, negative on error (SMX_EVALUE -1,
 *     SMX_ETYPE -2, SMX_ERUNTIME -3, SMX_ENUMERIC -4); smx_last_error()
 *     returns a thread-local message.  Data-dependent failures found inside
 *     kernels are reported through a device `status` word (bit flags) that the
 *     caller reads after synchronising.
 *   - float64 throughout the parity paths; the library is compiled with
 *     -ffp-contract=off so float64 expressions round like numpy's.
 */
#ifndef SCOPYON_HIP_H
#define SCOPYON_HIP_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#include "scopyon_hip_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- plumbing ----------------------------------------------------------- */
const char* smx_last_error(void);
int smx_version(void);
int smx_device_count(int* count /* host */);
/* Host evaluations of the device helpers (for CPU-side tests). */
void smx_philox4x32_10_host(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double smx_i1e_host(double y);                       /* scipy.special.i1e, _epifm.py:13,369 */
double smx_np_remainder_host(double a, double b);    /* numpy '%', sampling2.py:46-47 */
size_t smx_sat_index_host(int side, int period, int a, int b);

/* ---- PSF tables (_epifm.py:76-213) -------------------------------------- */
/* Born-Wolf radial profile for nz depths: out[z * nr + r], r = k * res.
 * Replaces PointSpreadingFunction.__get_born_wolf_distribution_old
 * (_epifm.py:180-213).  alpha = 2 pi NA / lambda, gamma = (2 pi / lambda)(NA/2)^2,
 * scale = alpha^2 / pi, all computed by the caller in float64.
 * j0_work: nr * 100 doubles. */
int smx_psf_radial_born_wolf(const double* depths, int nz, int nr, double res, double alpha,
                             double gamma, double scale, double* j0_work, double* out,
                             hipStream_t stream);
/* Gaussian radial profile (_epifm.py:128-134). */
int smx_psf_radial_gaussian(int nr, double res, double width, double* out, hipStream_t stream);
/* Summed-area tables of the (2 nr - 1)^2 cartesian tables
 * (radial_to_cartesian, _epifm.py:98-126) for nz profiles, stored with
 * smx_psf_sat_elems(nr, period) doubles per table. */
size_t smx_psf_sat_elems(int nr, int period);
size_t smx_psf_sat_workspace_bytes(int nr, int nz);
int smx_psf_sat_build(const double* radial, int nz, int nr, double res, int period,
                      double* workspace, double* sat, hipStream_t stream);

/* Radial x depth lookup-table PSF (smx_lut_t, scopyon_hip_types.h): per
 * depth slice z, the float32 pixel integrals K[qu][qv][du][dv]
 * (smx_lut_table_elems floats) of the cartesian kernel built from the
 * slice's float64 radial profile radial[z][k], k * r_max / (nb - 1), k < nb
 * (smx_psf_radial_born_wolf with res = r_max / (nb - 1); the cartesian
 * kernel of _epifm.py:98-126, the Born-Wolf profile of _epifm.py:180-213).
 * Each entry integrates the kernel over the exact pixel square by the
 * midpoint rule on cells of pixel_length / (q ns) (a float64 summed-area
 * table per slice, in `workspace`, processed in as many slices at a time as
 * it holds: smx_lut_build_workspace_bytes(lut, 1) per slice). */
size_t smx_lut_table_elems(const smx_lut_t* lut /* host */);
size_t smx_lut_build_workspace_bytes(const smx_lut_t* lut /* host */, int nz_batch);
int smx_lut_build(const double* radial, int nz, int nb, const smx_lut_t* lut /* host */,
                  double* workspace, size_t workspace_bytes, float* K, hipStream_t stream);
/* The splat's form of the tables (no reference counterpart): per slice and
 * cell (cu, cv), cu, cv < q, for each footprint pixel the four phases of
 * the bilinear taps around it, (cu, cv), (cu, cv + 1), (cu + 1, cv),
 * (cu + 1, cv + 1), as one 16-byte entry (smx_lut_cell_elems floats per
 * slice) -- one contiguous load per pixel instead of four. */
size_t smx_lut_cell_elems(const smx_lut_t* lut /* host */);
int smx_lut_cells(const float* K, int nz, const smx_lut_t* lut /* host */, float* cells,
                  hipStream_t stream);

/* ---- emitters and splat (_epifm.py:1267-1333, 215-282) ------------------ */
/* Per-emitter weight w (the reference's `normalization`) and depth-table key
 * for one sub-step of length unit_time.  Replaces the scalar part of
 * _EPIFMSimulator.__overlay_molecule_plane (_epifm.py:1280-1315); with
 * `budget` non-NULL also the photobleaching budget update (_epifm.py:1294-1306).
 * depth = particle column 0 (absolute); p_state may be NULL (all ones);
 * slot_map (may be NULL) maps depth-key index (0.., cutoff_slot) to table slot. */
int smx_emitter_weights(int n, const double* depth, const double* p_state,
                        const smx_weight_params_t* params /* host */, double unit_time,
                        double* budget, const int* budget_slot, const int* slot_map,
                        int n_slot_map, double* w_out, int* key_out, double* pstate_out,
                        hipStream_t stream);
/* Tiles of a frame (16 x 16 pixels) and the splat workspace size. */
int smx_splat_tiles(const smx_fov_t* fov /* host */, int reserved);
int smx_splat_tile_size(void);
size_t smx_splat_workspace_bytes(int n_emitters, int n_tiles, long long list_capacity);
/* Phase tables for pixel lengths of `sat_period` (integer >= 2) table cells:
 * K[slot][phi_u][phi_v][du][dv] = photons of footprint pixel (du, dv) for
 * boundary phases (phi_u, phi_v): the SAT integral times unit_area, 0 unless
 * positive; m = footprint pixels per axis.  Lets smx_splat_expected integrate
 * each (emitter, pixel) with one load. */
size_t smx_phase_table_elems(int sat_period, int m);
int smx_phase_table_build(const double* sat, size_t sat_stride, int nslots, int ntab,
                          int sat_period, int m, double unit_area, double* K,
                          hipStream_t stream);
/* float32 copy of a phase table (n elements) for the reduced-traffic splat. */
int smx_phase_table_to_float(const double* K, size_t n, float* K32, hipStream_t stream);
/* Folded float32 phase tables: the cartesian PSF table is symmetric under
 * reflection of either axis and under transposition, so pixel integrals of
 * phase ph equal those of ph' = (ntab - 2 - ph) mod P in reverse pixel order,
 * and (ph_u, ph_v) equals (ph_v, ph_u) transposed.  One np(0) x np(0) block per
 * unordered pair of canonical phases keeps every value in ~1/8 of the space
 * (2.5 MB per depth key at 160 nm pixels: small enough to stay in one XCD's
 * L2).  smx_phase_fold_host reports the canonical rank of phase ph, whether ph
 * is the reflected member, and its non-empty pixel count. */
size_t smx_phase_fold_elems(int ntab, int period);
int smx_phase_fold_host(int ntab, int period, int ph, int* rank /* host */,
                        int* reflected /* host */, int* npix /* host */);
int smx_phase_table_fold(const double* K, int nslots, int ntab, int period, int m, float* Kf,
                         hipStream_t stream);
/* Pixel integration of n emitters into nfov frames: replaces
 * PointSpreadingFunction.overlay_signal_ called per emitter
 * (_epifm.py:223-282 via _epifm.py:1261-1264).  u, v: absolute lateral
 * coordinates; w: weights; key: SAT slot per emitter; fov_id: FOV per emitter
 * (NULL: all in FOV 0).  sat: tables of sat_stride doubles each.
 * accumulate=1 adds into `out` (camera += expected); 0 overwrites.
 * K (may be NULL): phase tables for the same slots, k_format 0: float64 from
 * smx_phase_table_build (identical results, one load per emitter-pixel
 * instead of four); 1: its float32 copy (smx_phase_table_to_float); 2: its
 * folded float32 copy (smx_phase_table_fold).  With 1 and 2 each
 * contribution is rounded to float32.
 * status bit 1: a tile list held more than 8192 entries, too many to sort
 * (still correct, order not by emitter index; shorter lists are sorted in
 * the tile pass or, past its LDS capacity, after binning); bit 2: list
 * capacity too small (result
 * incomplete); bit 8: an emitter's table slot was not built (skipped). */
int smx_splat_expected(int n, const double* u, const double* v, const double* w, const int* key,
                       const int* fov_id, const smx_fov_t* fovs_dev, int nfov, int n_tiles,
                       const double* sat, size_t sat_stride, int n_slots, int ntab, int sat_period,
                       double res, double unit_area, long long list_capacity, int accumulate,
                       const void* K, int k_format, int m, double* out, void* workspace,
                       size_t workspace_bytes, int* status_dev, hipStream_t stream);
/* The two halves of smx_splat_expected, same arguments: smx_splat_bin bins
 * the emitters to tiles (and builds their table records) in `workspace`;
 * smx_splat_accumulate integrates the tiles from that workspace into `out`.
 * Lets a caller time or repeat the tile pass on its own. */
int smx_splat_bin(int n, const double* u, const double* v, const double* w, const int* key,
                  const int* fov_id, const smx_fov_t* fovs_dev, int nfov, int n_tiles,
                  const double* sat, size_t sat_stride, int n_slots, int ntab, int sat_period,
                  double res, double unit_area, long long list_capacity, int accumulate,
                  const void* K, int k_format, int m, double* out, void* workspace,
                  size_t workspace_bytes, int* status_dev, hipStream_t stream);
int smx_splat_accumulate(int n, const double* u, const double* v, const double* w,
                         const int* key, const int* fov_id, const smx_fov_t* fovs_dev, int nfov,
                         int n_tiles, const double* sat, size_t sat_stride, int n_slots, int ntab,
                         int sat_period, double res, double unit_area, long long list_capacity,
                         int accumulate, const void* K, int k_format, int m, double* out,
                         void* workspace, size_t workspace_bytes, int* status_dev,
                         hipStream_t stream);
/* The same three with the tile side as the last argument (16, 32 or 64 with
 * phase tables K, 32 without): the FOV descriptors' tiles_r / tiles_c and
 * n_tiles must be counted in tiles of that side.  The plain entry points use
 * 32 (smx_splat_tile_size()).  A tile of side 16 q is processed by q^2 waves,
 * one per 16 x 16 quadrant; with SMX_TILE_Q32 or-ed into a side of 32 or 64
 * (phase tables only) one wave takes each 32 x 32 quadrant instead (fewer
 * footprint pieces per emitter, more LDS per wave).  Binning and results do
 * not depend on the quadrant.  The batched path times the shapes on its own
 * workload and keeps the fastest.  No reference counterpart (the reference
 * loops over emitters on the whole frame, _epifm.py:215-282). */
#define SMX_TILE_Q32 0x100
/* Or-ed into the tile side: this workspace was binned before by a call with
 * the same FOV layout, tile side, emitter count and list capacity, so its
 * tile counters are back at zero (the fill counts them down) and its tile ->
 * FOV map is current; the binning then skips their initialisation (one
 * launch).  Only with fewer than 65,536 tiles (the one-workgroup counter
 * scan); ignored otherwise.  No reference counterpart. */
#define SMX_TILE_WARM 0x200
int smx_splat_expected_ts(int n, const double* u, const double* v, const double* w,
                          const int* key, const int* fov_id, const smx_fov_t* fovs_dev, int nfov,
                          int n_tiles, const double* sat, size_t sat_stride, int n_slots,
                          int ntab, int sat_period, double res, double unit_area,
                          long long list_capacity, int accumulate, const void* K, int k_format,
                          int m, double* out, void* workspace, size_t workspace_bytes,
                          int* status_dev, hipStream_t stream, int tile_side);
int smx_splat_bin_ts(int n, const double* u, const double* v, const double* w, const int* key,
                     const int* fov_id, const smx_fov_t* fovs_dev, int nfov, int n_tiles,
                     const double* sat, size_t sat_stride, int n_slots, int ntab, int sat_period,
                     double res, double unit_area, long long list_capacity, int accumulate,
                     const void* K, int k_format, int m, double* out, void* workspace,
                     size_t workspace_bytes, int* status_dev, hipStream_t stream, int tile_side);
int smx_splat_accumulate_ts(int n, const double* u, const double* v, const double* w,
                            const int* key, const int* fov_id, const smx_fov_t* fovs_dev,
                            int nfov, int n_tiles, const double* sat, size_t sat_stride,
                            int n_slots, int ntab, int sat_period, double res, double unit_area,
                            long long list_capacity, int accumulate, const void* K, int k_format,
                            int m, double* out, void* workspace, size_t workspace_bytes,
                            int* status_dev, hipStream_t stream, int tile_side);

/* Splat and EMCCD detector in one pass over the tiles (phase tables K
 * required, accumulate = 0): each tile's photons go from the tile pass's
 * on-chip accumulators through the EMCCD register, readout noise and ADC
 * (_epifm.py:1436-1469) of smx_emccd_philox -- the same Philox draws per
 * pixel, so adc_out equals smx_splat_expected_ts followed by
 * smx_emccd_philox(pixel count, out, gain_map, offset_map, params, seed,
 * frame, pix_base, frame_dev, E_out, adc_out, adc_is_float32) bit for bit --
 * without writing the photon image to memory and reading it back.  `out`
 * (the float64 photons) may be NULL; when given it is written too.
 * smx_splat_detect_accumulate_ts is the tile pass alone on an already binned
 * workspace (the second half, like smx_splat_accumulate_ts).  No reference
 * counterpart as one call (the reference forms the image, then runs the
 * detector, _epifm.py:1199, 1220). */
int smx_splat_detect_ts(int n, const double* u, const double* v, const double* w, const int* key,
                        const int* fov_id, const smx_fov_t* fovs_dev, int nfov, int n_tiles,
                        const double* sat, size_t sat_stride, int n_slots, int ntab,
                        int sat_period, double res, double unit_area, long long list_capacity,
                        int accumulate, const void* K, int k_format, int m, double* out,
                        void* workspace, size_t workspace_bytes, int* status_dev,
                        hipStream_t stream, const double* gain_map, const double* offset_map,
                        const smx_detector_params_t* params /* host */, uint64_t seed,
                        uint32_t frame, uint32_t pix_base, const uint32_t* frame_dev,
                        double* E_out, void* adc_out, int adc_is_float32, int tile_side);
int smx_splat_detect_accumulate_ts(int n, const double* u, const double* v, const double* w,
                                   const int* key, const int* fov_id, const smx_fov_t* fovs_dev,
                                   int nfov, int n_tiles, const double* sat, size_t sat_stride,
                                   int n_slots, int ntab, int sat_period, double res,
                                   double unit_area, long long list_capacity, int accumulate,
                                   const void* K, int k_format, int m, double* out,
                                   void* workspace, size_t workspace_bytes, int* status_dev,
                                   hipStream_t stream, const double* gain_map,
                                   const double* offset_map,
                                   const smx_detector_params_t* params /* host */, uint64_t seed,
                                   uint32_t frame, uint32_t pix_base, const uint32_t* frame_dev,
                                   double* E_out, void* adc_out, int adc_is_float32,
                                   int tile_side);

/* Pixel integration in LUT mode: the overlay of _epifm.py:215-282 with the
 * pixel integrals of lut_tables (the cell form, smx_lut_cells; key[i] =
 * depth slice of emitter i) instead of 1 nm cell sums.  Emitter i at a = (W/2 + x) / pl
 * pixels from the frame edge takes the bilinear combination of the four
 * tabulated phases around a q.  Same tiles, binning, workspace
 * (smx_splat_workspace_bytes), list order and status bits as
 * smx_splat_expected_ts; stages: 1 bin, 2 tile pass on a binned workspace,
 * 3 both.  smx_splat_lut_detect_ts fuses the EMCCD detector into the tile
 * pass like smx_splat_detect_ts (accumulate = 0; out may be NULL).  No
 * reference counterpart (a deliberate deviation: the reference keeps one
 * 1 nm table per 1 nm depth key). */
int smx_splat_lut_ts(int n, const double* u, const double* v, const double* w, const int* key,
                     const int* fov_id, const smx_fov_t* fovs_dev, int nfov, int n_tiles,
                     const float* lut_tables, const smx_lut_t* lut /* host */,
                     long long list_capacity, int accumulate, double* out, void* workspace,
                     size_t workspace_bytes, int* status_dev, hipStream_t stream, int stages,
                     int tile_side);
int smx_splat_lut_detect_ts(int n, const double* u, const double* v, const double* w,
                            const int* key, const int* fov_id, const smx_fov_t* fovs_dev, int nfov,
                            int n_tiles, const float* lut_tables, const smx_lut_t* lut /* host */,
                            long long list_capacity, int accumulate, double* out, void* workspace,
                            size_t workspace_bytes, int* status_dev, hipStream_t stream,
                            const double* gain_map, const double* offset_map,
                            const smx_detector_params_t* params /* host */, uint64_t seed,
                            uint32_t frame, uint32_t pix_base, const uint32_t* frame_dev,
                            double* E_out, void* adc_out, int adc_is_float32, int stages,
                            int tile_side);

/* Allocation-free form of smx_splat_expected_ts, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  Same
 * arguments without the workspace: the binning workspace lives in a fixed
 * device array of smx_splat_static_scratch_bytes() (64 MiB) that each GPU the
 * library runs on holds until the process ends; a call whose workspace
 * (smx_splat_workspace_bytes(n, n_tiles, list_capacity)) exceeds it returns
 * SMX_EVALUE.  Calls that overlap on one GPU share that array and race:
 * order them on one stream.  Nothing here calls hipMalloc or hipFree. */
size_t smx_splat_static_scratch_bytes(void);
int smx_splat_expected_static(int n, const double* u, const double* v, const double* w,
                              const int* key, const int* fov_id, const smx_fov_t* fovs_dev,
                              int nfov, int n_tiles, const double* sat, size_t sat_stride,
                              int n_slots, int ntab, int sat_period, double res, double unit_area,
                              long long list_capacity, int accumulate, const void* K, int k_format,
                              int m, double* out, int* status_dev, hipStream_t stream,
                              int tile_side);

/* ---- detector (_epifm.py:355-416, 1430-1484) ---------------------------- */
/* E = QE * (photons + background) (_epifm.py:1436-1441); flag_out (may be
 * NULL) = E > 0, the pixels that consume a uniform in the EMCCD draw. */
int smx_detector_expectation(int n, const double* photons,
                             const smx_detector_params_t* params /* host */, double* E_out,
                             int* flag_out, hipStream_t stream);
size_t smx_scan_workspace_bytes(int n);
int smx_exclusive_scan_i32(int n, const int* in, int* out, void* workspace, size_t bytes,
                           hipStream_t stream);
/* The pmf of one expectation value shared by many pixels (the background-
 * only ones): T_out (device, smx_emccd_common_elems() doubles) receives the
 * total mass and the partial sums smx_emccd_parity searches. */
size_t smx_emccd_common_elems(void);
int smx_emccd_common_total(double E, int emgain, double* T_out, hipStream_t stream);
/* EMCCD.get_signal + ADC with caller-drawn random numbers
 * (_epifm.py:393-416 and 1472-1484): noise[i] = readout noise (NULL: 0),
 * uniforms[uidx[i]] = the uniform RandomState.choice draws for pixel i.
 * Pixels with E == E_common reuse T_common (smx_emccd_common_total's output
 * for E_common; may be NULL).  The draw: the support's terms in chunks of
 * 512 (8 consecutive terms per lane), the chunk where the cumulative mass
 * passes u T found from the chunk sums, then the lane and the index inside
 * it (RandomState.choice's searchsorted(cdf, u, 'right') on the normalised
 * cumulative sum, equal unless u falls within rounding of a cdf step).  status bit 4:
 * the pmf overflowed for some pixel (the reference raises ValueError). */
int smx_emccd_parity(int n, const double* E, const double* noise, const double* uniforms,
                     const int* uidx, const double* gain_map, const double* offset_map,
                     const smx_detector_params_t* params /* host */, double E_common,
                     const double* T_common, double* adc_out, double* signal_out,
                     int* status_dev, hipStream_t stream);
/* EMCCD with Philox4x32-10 randomness: Poisson(E) -> Gamma(n, emgain) signal,
 * Gaussian readout noise, ADC.  Fuses _epifm.py:1436-1469 in float32
 * arithmetic (a statistical path; the parity path is smx_emccd_parity).  The Philox frame
 * counter is `frame` plus *frame_dev when frame_dev is not NULL. */
int smx_emccd_philox(int n, const double* photons, const double* gain_map,
                     const double* offset_map, const smx_detector_params_t* params /* host */,
                     uint64_t seed, uint32_t frame, uint32_t pix_base,
                     const uint32_t* frame_dev, double* E_out, void* adc_out,
                     int adc_is_float32, hipStream_t stream);

/* CCD and CMOS with caller-drawn numbers (_epifm.py:329-353, 418-433,
 * 1443-1469): adc = ADC(signal + noise) with signal the caller's
 * RandomState.poisson photo-electrons and noise either noise_in[i] (CCD
 * normal draws; NULL: none) or, when table_n > 0, the CMOS readout value
 * table_values[searchsorted(table_cdf, noise_in[i], 'right')] -- the draw of
 * RandomState.choice(values, p) given its uniform noise_in[i]. */
int smx_adc_parity(int n, const double* signal, const double* noise_in, const double* table_values,
                   const double* table_cdf, int table_n, const double* gain_map,
                   const double* offset_map, const smx_detector_params_t* params /* host */,
                   double* adc_out, hipStream_t stream);
/* CCD (table_n == 0: Gaussian readout noise of params->readout_noise) or CMOS
 * (readout noise drawn from the value table) with Philox4x32-10 randomness:
 * Poisson(E) photo-electrons + noise, ADC; float32 arithmetic. */
int smx_photon_detector_philox(int n, const double* photons, const double* gain_map,
                               const double* offset_map,
                               const smx_detector_params_t* params /* host */,
                               const double* table_values, const double* table_cdf, int table_n,
                               uint64_t seed, uint32_t frame, uint32_t pix_base,
                               const uint32_t* frame_dev, double* E_out, void* adc_out,
                               int adc_is_float32, hipStream_t stream);

/* ---- separable Gaussian model: forward, Poisson NLL, backward ----------- */
/* Frames reuse the smx_fov_t fields nw, nh, tiles_r, tiles_c, pix_off and
 * tile_off (the length fields are unused: coordinates are pixels).  Emitter e belongs to
 * frame frame_id[e] (NULL: frame 0).  The model is documented at
 * smx_gauss_params_t; the Gaussian profile is _epifm.py:128-134 and the depth
 * factor the TIRF amplitude of _epifm.py:1289.  Float32 throughout; the
 * per-frame loss and the PSF-width gradient are reduced in float64. */
size_t smx_gauss_workspace_bytes(int n, int n_tiles, long long capacity);
/* mu = background + sum of emitter contributions (every pixel of every frame
 * is written).  capacity: tile-list entries (each emitter enters at most four
 * 32 x 32 tiles).  status bit 1: a tile list held more than 8192 entries,
 * too many to sort (result correct, summation order not by emitter index);
 * bit 2: capacity too
 * small (result incomplete). */
int smx_gauss_forward(int n, const float* x, const float* y, const float* z, const float* I,
                      const int* frame_id, const smx_fov_t* frames_dev, int nframes, int n_tiles,
                      const smx_gauss_params_t* params /* host */, long long capacity, float* mu,
                      void* workspace, size_t workspace_bytes, int* status_dev,
                      hipStream_t stream);
/* The same image on the matrix cores, where the PSF is separable: per 32 x 32
 * tile, mu = background + A B with A[i][e] = I_e a(z_e) G(i; x_e) and
 * B[e][j] = G(j; y_e) over the tile's emitters (zero outside each window),
 * the dense contraction Gx^T diag(I) Gy.  dtype 0: float32 operands
 * (v_mfma_f32_32x32x2_f32), 1: float16, 2: bfloat16 (v_mfma_f32_32x32x16_*),
 * float32 accumulation.  z may be NULL (no depth factor).  Same workspace and
 * status bits as smx_gauss_forward. */
int smx_sep_gemm(int n, const float* x, const float* y, const float* z, const float* I,
                 const int* frame_id, const smx_fov_t* frames_dev, int nframes, int n_tiles,
                 const smx_gauss_params_t* params /* host */, int dtype, long long capacity,
                 float* mu, void* workspace, size_t workspace_bytes, int* status_dev,
                 hipStream_t stream);
/* smx_sep_gemm with the Poisson NLL fused into the tile pass: loss[f] = sum
 * over frame f of mu - k log(mu) (float64 terms as smx_gauss_poisson_nll,
 * summed per tile, then over the frame's tiles in order) and grad_mu (may be
 * NULL) = 1 - k / mu, from the tile's accumulators; mu (may be NULL) is
 * written only when given.  Same workspace as smx_sep_gemm. */
int smx_sep_gemm_nll(int n, const float* x, const float* y, const float* z, const float* I,
                     const int* frame_id, const smx_fov_t* frames_dev, int nframes, int n_tiles,
                     const smx_gauss_params_t* params /* host */, int dtype, long long capacity,
                     float* mu, const float* k, float* grad_mu, double* loss, void* workspace,
                     size_t workspace_bytes, int* status_dev, hipStream_t stream);
/* loss[f] = sum over frame f of mu - k log(mu) (float64); grad_mu (may be NULL)
 * = 1 - k / mu.  workspace: smx_gauss_nll_workspace_bytes(nframes, max pixels
 * of any frame). */
size_t smx_gauss_nll_workspace_bytes(int nframes, int max_frame_pixels);
int smx_gauss_poisson_nll(const float* mu, const float* k, const smx_fov_t* frames_dev,
                          int nframes, int max_frame_pixels, double* loss, float* grad_mu,
                          void* workspace, size_t workspace_bytes, hipStream_t stream);
/* Hand-written backward of smx_gauss_forward: given grad_mu = dL/dmu, writes
 * dL/dx, dL/dy, dL/dz, dL/dI per emitter (any output may be NULL) and, when
 * gsigma (device scalar) is not NULL, dL/dsigma summed over all emitters in a
 * fixed order (float64).  The window factors G, dG/dx, dG/dsigma come from a
 * table by sub-pixel phase built in float64 for params->sigma (1025 phases,
 * linear interpolation).  workspace: smx_gauss_backward_workspace_bytes(n). */
size_t smx_gauss_backward_workspace_bytes(int n);
int smx_gauss_backward(int n, const float* x, const float* y, const float* z, const float* I,
                       const int* frame_id, const smx_fov_t* frames_dev, int nframes,
                       const smx_gauss_params_t* params /* host */, const float* grad_mu,
                       float* gx, float* gy, float* gz, float* gI, double* gsigma,
                       void* workspace, size_t workspace_bytes, hipStream_t stream);
/* The same gradients from the binning of the forward call that produced
 * mu (smx_gauss_forward, smx_sep_gemm or smx_sep_gemm_nll with the same n,
 * emitters, frames, n_tiles, params and capacity, whose workspace is passed
 * here untouched, on the same stream): each tile stages its window of
 * grad_mu in LDS and reduces the emitters whose window centre (clamped into
 * the frame) it holds; dL/dsigma is summed per tile in list order, then over
 * the tiles.  gx, gy, gz, gI as smx_gauss_backward (bit-identical values);
 * gsigma the same sum in another fixed order.  status bit 16: the workspace
 * holds the binning of another sigma, radius, n or n_tiles (results
 * invalid). */
int smx_gauss_backward_binned(int n, const float* x, const float* y, const float* z,
                              const float* I, const int* frame_id, const smx_fov_t* frames_dev,
                              int nframes, int n_tiles,
                              const smx_gauss_params_t* params /* host */, long long capacity,
                              const float* grad_mu, float* gx, float* gy, float* gz, float* gI,
                              double* gsigma, void* workspace, size_t workspace_bytes,
                              int* status_dev, hipStream_t stream);

/* ---- Brownian sampling (sampling2.py:33-68, sampling.py:88-119) --------- */
/* Whole trajectories from caller-drawn normals xi[t][i][d] (and transition
 * uniforms trans_u[t][i]): x += scale[t][state][d] * xi; boundary 0 none,
 * 1 periodic wrap with numpy '%' semantics, 2 reflecting walls (device
 * lower/upper); optional state transitions
 * searchsorted(pacc[t][state], u, 'left').  xs[(t+1)][i][d], states[t][i]. */
int smx_brownian_predrawn(int n, int ndim, int nsteps, int nstates, const double* x0,
                          const int* state0, const double* scale, const double* xi, int boundary,
                          const double* lower, const double* upper, const double* pacc,
                          const double* trans_u, double* xs, int* states, hipStream_t stream);
/* One Philox step on structure-of-arrays coordinates: x += sqrt(2 D dt) N(0,1);
 * boundary 0 none, 1 periodic, 2 reflecting.  lower/upper: host, ndim each;
 * or, with fov_box (device, 6 doubles per FOV: lo[3], hi[3]) and fov_id,
 * per-FOV walls.  Extension of sampling2.py:33-50 (the reference has only
 * the periodic wrap). */
int smx_brownian_philox(int n, int ndim, double* x0, double* x1, double* x2, const double* D,
                        double D_scalar, double dt, int boundary, const double* lower,
                        const double* upper, uint64_t seed, uint32_t step, uint32_t id_base,
                        const int* fov_id, int nfov, const double* fov_box,
                        const uint32_t* step_dev, hipStream_t stream);
/* One Philox step with per-state, per-axis diffusion and state transitions
 * (sampling2.py:33-68; per-axis D as in sampling.py:115-118):
 * x_d += sqrt(2 D_tab[state][d] dt) N(0, 1) (device D_tab, nstates x ndim),
 * boundary as smx_brownian_philox (lower/upper host), then, when pacc
 * (device, nstates x nstates cumulative transition rows) is given,
 * state = searchsorted(pacc[state], u, 'left') with u a Philox uniform of the
 * particle.  state (device int32) may be NULL for one state. */
int smx_brownian_philox_states(int n, int ndim, double* x0, double* x1, double* x2, int* state,
                               int nstates, const double* D_tab, double dt, int boundary,
                               const double* lower, const double* upper, const double* pacc,
                               uint64_t seed, uint32_t step, uint32_t id_base,
                               hipStream_t stream);
/* counter[0] += by on the device (advances Philox counters inside graphs). */
int smx_counter_add(uint32_t* counter, uint32_t by, hipStream_t stream);
/* Uniform initial positions (sampling2.py:15-31) from Philox; lower/upper host. */
int smx_uniform_init(int n, int ndim, double* x0, double* x1, double* x2, const double* lower,
                     const double* upper, uint64_t seed, uint32_t seq, uint32_t id_base,
                     hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* SCOPYON_HIP_H */
