"""ctypes binding of libscopyon_hip.so (the C-ABI declared in include/scopyon_hip.h).

The library is loaded on first use.  There is no CPU fallback: if the shared
library is missing or cannot be loaded every compute entry point raises
``RuntimeError`` naming the file.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libscopyon_hip.so')

c_int, c_double, c_size_t, c_void_p = ctypes.c_int, ctypes.c_double, ctypes.c_size_t, ctypes.c_void_p
c_longlong, c_uint32, c_uint64 = ctypes.c_longlong, ctypes.c_uint32, ctypes.c_uint64
P = c_void_p   # device or host pointer passed as an integer address

SMX_OK, SMX_EVALUE, SMX_ETYPE, SMX_ERUNTIME, SMX_ENUMERIC = 0, -1, -2, -3, -4


class FovT(ctypes.Structure):
    """smx_fov_t (include/scopyon_hip_types.h)."""
    _fields_ = [('nw', ctypes.c_int32), ('nh', ctypes.c_int32),
                ('tiles_r', ctypes.c_int32), ('tiles_c', ctypes.c_int32),
                ('pix_off', ctypes.c_int64), ('tile_off', ctypes.c_int32),
                ('reserved', ctypes.c_int32), ('pixel_length', c_double),
                ('half_w', c_double), ('half_h', c_double), ('p0u', c_double), ('p0v', c_double)]


class WeightParamsT(ctypes.Structure):
    """smx_weight_params_t."""
    _fields_ = [('amplitude', c_double), ('penetration_depth', c_double), ('x_sec', c_double),
                ('quantum_yield', c_double), ('one_minus_10_pow_minus_a', c_double),
                ('spectral_psf_norm', c_double), ('four_pi', c_double), ('focal_depth', c_double),
                ('depth_cutoff', c_double), ('res', c_double), ('cutoff_slot', ctypes.c_int32),
                ('reserved', ctypes.c_int32)]


class GaussParamsT(ctypes.Structure):
    """smx_gauss_params_t."""
    _fields_ = [('sigma', ctypes.c_float), ('inv_depth', ctypes.c_float),
                ('background', ctypes.c_float), ('radius', ctypes.c_int32)]


class LutT(ctypes.Structure):
    """smx_lut_t."""
    _fields_ = [('q', ctypes.c_int32), ('m', ctypes.c_int32), ('c', ctypes.c_int32),
                ('nslices', ctypes.c_int32), ('ns', ctypes.c_int32), ('reserved', ctypes.c_int32),
                ('pixel_length', c_double), ('r_max', c_double)]


class DetectorParamsT(ctypes.Structure):
    """smx_detector_params_t."""
    _fields_ = [('qe', c_double), ('background', c_double), ('background_switch', ctypes.c_int32),
                ('emgain', ctypes.c_int32), ('readout_noise', c_double), ('fullwell', c_double),
                ('adc_max', c_double), ('offset', c_double), ('gain', c_double)]


# name -> (restype, argtypes)
SIGNATURES = {
    'smx_last_error': (ctypes.c_char_p, []),
    'smx_version': (c_int, []),
    'smx_device_count': (c_int, [P]),
    'smx_philox4x32_10_host': (None, [P, P, P]),
    'smx_i1e_host': (c_double, [c_double]),
    'smx_np_remainder_host': (c_double, [c_double, c_double]),
    'smx_sat_index_host': (c_size_t, [c_int, c_int, c_int, c_int]),
    'smx_psf_radial_born_wolf': (c_int, [P, c_int, c_int, c_double, c_double, c_double, c_double,
                                         P, P, P]),
    'smx_psf_radial_gaussian': (c_int, [c_int, c_double, c_double, P, P]),
    'smx_psf_sat_elems': (c_size_t, [c_int, c_int]),
    'smx_psf_sat_workspace_bytes': (c_size_t, [c_int, c_int]),
    'smx_psf_sat_build': (c_int, [P, c_int, c_int, c_double, c_int, P, P, P]),
    'smx_emitter_weights': (c_int, [c_int, P, P, ctypes.POINTER(WeightParamsT), c_double, P, P,
                                    P, c_int, P, P, P, P]),
    'smx_splat_tiles': (c_int, [ctypes.POINTER(FovT), c_int]),
    'smx_splat_tile_size': (c_int, []),
    'smx_splat_workspace_bytes': (c_size_t, [c_int, c_int, c_longlong]),
    'smx_phase_table_elems': (c_size_t, [c_int, c_int]),
    'smx_phase_table_build': (c_int, [P, c_size_t, c_int, c_int, c_int, c_int, c_double, P, P]),
    'smx_splat_expected': (c_int, [c_int, P, P, P, P, P, P, c_int, c_int, P, c_size_t, c_int,
                                   c_int, c_int, c_double, c_double, c_longlong, c_int, P, c_int,
                                   c_int, P, P, c_size_t, P, P]),
    'smx_phase_table_to_float': (c_int, [P, c_size_t, P, P]),

    'smx_phase_fold_elems': (c_size_t, [c_int, c_int]),
    'smx_phase_fold_host': (c_int, [c_int, c_int, c_int, P, P, P]),
    'smx_phase_table_fold': (c_int, [P, c_int, c_int, c_int, c_int, P, P]),
    'smx_detector_expectation': (c_int, [c_int, P, ctypes.POINTER(DetectorParamsT), P, P, P]),
    'smx_scan_workspace_bytes': (c_size_t, [c_int]),
    'smx_exclusive_scan_i32': (c_int, [c_int, P, P, P, c_size_t, P]),
    'smx_emccd_common_elems': (c_size_t, []),
    'smx_emccd_common_total': (c_int, [c_double, c_int, P, P]),
    'smx_emccd_parity': (c_int, [c_int, P, P, P, P, P, P, ctypes.POINTER(DetectorParamsT),
                                 c_double, P, P, P, P, P]),
    'smx_emccd_philox': (c_int, [c_int, P, P, P, ctypes.POINTER(DetectorParamsT), c_uint64,
                                 c_uint32, c_uint32, P, P, P, c_int, P]),
    'smx_brownian_predrawn': (c_int, [c_int, c_int, c_int, c_int, P, P, P, P, c_int, P, P, P, P,
                                      P, P, P]),
    'smx_brownian_philox': (c_int, [c_int, c_int, P, P, P, P, c_double, c_double, c_int, P, P,
                                    c_uint64, c_uint32, c_uint32, P, c_int, P, P, P]),
    'smx_brownian_philox_states': (c_int, [c_int, c_int, P, P, P, P, c_int, P, c_double, c_int,
                                           P, P, P, c_uint64, c_uint32, c_uint32, P]),
    'smx_counter_add': (c_int, [P, c_uint32, P]),
    'smx_uniform_init': (c_int, [c_int, c_int, P, P, P, P, P, c_uint64, c_uint32, c_uint32, P]),
}

_lib = None


SIGNATURES.update({
    'smx_adc_parity': (c_int, [c_int, P, P, P, P, c_int, P, P, ctypes.POINTER(DetectorParamsT),
                               P, P]),
    'smx_photon_detector_philox': (c_int, [c_int, P, P, P, ctypes.POINTER(DetectorParamsT), P, P,
                                           c_int, c_uint64, c_uint32, c_uint32, P, P, P, c_int,
                                           P]),
    'smx_gauss_workspace_bytes': (c_size_t, [c_int, c_int, c_longlong]),
    'smx_gauss_forward': (c_int, [c_int, P, P, P, P, P, P, c_int, c_int,
                                  ctypes.POINTER(GaussParamsT), c_longlong, P, P, c_size_t, P, P]),
    'smx_sep_gemm': (c_int, [c_int, P, P, P, P, P, P, c_int, c_int, ctypes.POINTER(GaussParamsT),
                             c_int, c_longlong, P, P, c_size_t, P, P]),
    'smx_sep_gemm_nll': (c_int, [c_int, P, P, P, P, P, P, c_int, c_int,
                                 ctypes.POINTER(GaussParamsT), c_int, c_longlong, P, P, P, P, P,
                                 c_size_t, P, P]),
    'smx_gauss_backward_workspace_bytes': (c_size_t, [c_int]),
    'smx_gauss_nll_workspace_bytes': (c_size_t, [c_int, c_int]),
    'smx_gauss_poisson_nll': (c_int, [P, P, P, c_int, c_int, P, P, P, c_size_t, P]),
    'smx_gauss_backward': (c_int, [c_int, P, P, P, P, P, P, c_int, ctypes.POINTER(GaussParamsT),
                                   P, P, P, P, P, P, P, c_size_t, P]),
    'smx_gauss_backward_binned': (c_int, [c_int, P, P, P, P, P, P, c_int, c_int,
                                          ctypes.POINTER(GaussParamsT), c_longlong, P, P, P, P,
                                          P, P, P, c_size_t, P, P]),
})
SIGNATURES['smx_splat_bin'] = SIGNATURES['smx_splat_expected']
SIGNATURES['smx_splat_accumulate'] = SIGNATURES['smx_splat_expected']
for _name in ('smx_splat_expected', 'smx_splat_bin', 'smx_splat_accumulate'):
    SIGNATURES[_name + '_ts'] = (c_int, SIGNATURES[_name][1] + [c_int])
# smx_splat_expected + (gain_map, offset_map, params, seed, frame, pix_base,
# frame_dev, E_out, adc_out, adc_is_float32, tile_side)
for _name in ('smx_splat_detect_ts', 'smx_splat_detect_accumulate_ts'):
    SIGNATURES[_name] = (c_int, SIGNATURES['smx_splat_expected'][1] + [
        P, P, ctypes.POINTER(DetectorParamsT), c_uint64, c_uint32, c_uint32, P, P, P, c_int,
        c_int])
_LUT = ctypes.POINTER(LutT)
SIGNATURES['smx_lut_table_elems'] = (c_size_t, [_LUT])
SIGNATURES['smx_lut_build_workspace_bytes'] = (c_size_t, [_LUT, c_int])
SIGNATURES['smx_lut_build'] = (c_int, [P, c_int, c_int, _LUT, P, c_size_t, P, P])
SIGNATURES['smx_lut_cell_elems'] = (c_size_t, [_LUT])
SIGNATURES['smx_lut_cells'] = (c_int, [P, c_int, _LUT, P, P])
_LUT_SPLAT = [c_int, P, P, P, P, P, P, c_int, c_int, P, _LUT, c_longlong, c_int, P, P, c_size_t,
              P, P]
SIGNATURES['smx_splat_lut_ts'] = (c_int, _LUT_SPLAT + [c_int, c_int])
SIGNATURES['smx_splat_lut_detect_ts'] = (c_int, _LUT_SPLAT + [
    P, P, ctypes.POINTER(DetectorParamsT), c_uint64, c_uint32, c_uint32, P, P, P, c_int, c_int,
    c_int])
# smx_splat_expected_ts without (workspace, workspace_bytes)
SIGNATURES['smx_splat_static_scratch_bytes'] = (c_size_t, [])
SIGNATURES['smx_splat_expected_static'] = (
    c_int, SIGNATURES['smx_splat_expected'][1][:22] + SIGNATURES['smx_splat_expected'][1][24:] +
    [c_int])


def lib():
    """Load (once) and return the ctypes library; raise if it is unavailable."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                'scopyon_amd: HIP library not built ({}); run __graft_entry__.build() or '
                '`make -C scopyon_amd/csrc`'.format(LIB_PATH))
        try:
            handle = ctypes.CDLL(LIB_PATH)
        except OSError as e:
            raise RuntimeError('scopyon_amd: cannot load {}: {}'.format(LIB_PATH, e))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def host_doubles(values, n=None):
    """A ctypes double array (host memory) for `/* host */` pointer arguments."""
    vals = [float(v) for v in values]
    if n is not None:
        vals = (vals + [0.0] * n)[:n]
    return (ctypes.c_double * max(len(vals), 1))(*vals)


def check(rc, what=''):
    """Map a C-ABI status code to the reference's Python exception types."""
    if rc == SMX_OK:
        return
    msg = lib().smx_last_error().decode(errors='replace') or what
    if rc == SMX_EVALUE or rc == SMX_ENUMERIC:
        raise ValueError(msg)
    if rc == SMX_ETYPE:
        raise TypeError(msg)
    raise RuntimeError(msg)


def call(name, *args):
    """Call a status-returning entry point and raise on failure."""
    check(getattr(lib(), name)(*args), name)
