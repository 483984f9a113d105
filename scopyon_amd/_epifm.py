"""EPIFM/TIRF image formation on the GPU.

Host side (this module): configuration scalars (``EPIFMConfigs``,
src/scopyon/_epifm.py:757-996), the frame/sub-step bookkeeping of
``output_frame`` (_epifm.py:1121-1225) and random-number plumbing.  Device side
(libscopyon_hip.so): emitter weights, PSF tables, pixel integration, EMCCD and
ADC.  There is no CPU fallback for any of the device steps.

Random numbers follow the caller's generator:
  * ``numpy.random.RandomState``: reference stream.  Every draw the reference
    makes (readout noise, the per-pixel ``choice`` uniforms, bleaching
    budgets, fixed-pattern offsets) is taken from it in the reference's order,
    so the ADC output matches the reference for the same seed.
  * ``scopyon_amd.PhiloxRNG``: counter-based Philox4x32-10 on the device; no
    host draws, statistically equivalent.
"""
import copy
import os
import math
import warnings

import numpy
import torch

from . import _device, _lib, io
from ._psf import PSFTables, RES, depth_keys
from .constants import N_A, hc
from .rng import PhiloxRNG

from logging import getLogger
_log = getLogger(__name__)


# --------------------------------------------------------------------------
# Host-side physics scalars
# --------------------------------------------------------------------------

def calculate_efficiency(data, wave_length):
    """_epifm.py:983-996: catalog rows on integer nm aligned to the 1 nm grid."""
    data = numpy.array(data, dtype=float)
    data = data[data[:, 0] % 1 == 0, :]
    eff = numpy.zeros(len(wave_length))
    grid = numpy.round(wave_length / 1e-9).astype(int)
    eff[numpy.isin(grid, data[:, 0])] = data[numpy.isin(data[:, 0], grid), 1]
    return eff.tolist()


def snells_law(flux_density, wave_length, angle):
    """_epifm.py:1362-1428 -> (amplitude, penetration depth)."""
    N_0 = flux_density / (hc / wave_length)
    sin, cos = numpy.sin(angle), numpy.cos(angle)
    sin2, cos2 = sin ** 2, cos ** 2
    n_1, n_2 = 1.46, 1.384
    r2 = (n_2 / n_1) ** 2
    if sin2 / r2 < 1:
        return N_0, numpy.inf
    A2_x = N_0 * (4 * cos2 * (sin2 - r2) / (r2 ** 2 * cos2 + sin2 - r2))
    A2_y = N_0 * (4 * cos2 / (1 - r2))
    A2_z = N_0 * (4 * cos2 * sin2 / (r2 ** 2 * cos2 + sin2 - r2))
    amplitude = ((A2_x + A2_z) + A2_y) / 2
    depth = wave_length / (4.0 * numpy.pi * numpy.sqrt(n_1 ** 2 * sin2 - n_2 ** 2))
    return amplitude, depth


def emit_constants(abs_coeff, radius):
    """The unit_time-independent factors of get_emit_photons (_epifm.py:1343-1360)."""
    x_sec = numpy.log(10) * abs_coeff * 0.1 / N_A
    volume = (4.0 / 3.0) * numpy.pi * numpy.power(radius, 3)
    A = (abs_coeff * 0.1 / N_A) * (1.0 / volume) * (2.0 * radius)
    return x_sec, 1.0 - numpy.power(10.0, -A)


def get_emit_photons(amplitude, unit_time, abs_coeff, QY, radius):
    """_epifm.py:1343-1360 (host evaluation, e.g. for the bleaching budget)."""
    x_sec, one_minus = emit_constants(abs_coeff, radius)
    return QY * (amplitude * x_sec * unit_time) * one_minus


class EPIFMConfigs:
    """Scalars of one ``default``-like config sub-tree (_epifm.py:757-996)."""

    def __init__(self, config, rng=None):
        self.initialize(config, rng=rng)

    def initialize(self, config, rng=None):
        if config.type.lower() != 'epifm':
            raise ValueError("An invalid type [{}] was given. 'epifm' is required.".format(config.type))
        if rng is None:
            warnings.warn('A random number generator [rng] is not given.')
            rng = numpy.random.RandomState()
        self.set_fluorophore(**config.fluorophore)
        self.shutter = {k: getattr(config.shutter, k) for k in config.shutter.yaml}
        ls = config.light_source
        self.source_wavelength = ls.wave_length
        self.source_flux_density = ls.flux_density
        self.source_angle = ls.angle
        self.dichroic_switch = config.dichroic_mirror.switch
        self.emission_switch = config.emission_filter.switch
        # _epifm.py:1308-1313 scales fluoem_norm in place by 0.01 * efficiency.
        # An efficiency read from a catalog is a Python list (_epifm.py:996),
        # so the reference raises TypeError at the first emitter; without a
        # type the efficiency is zeros and every weight becomes 0 for good.
        self.filter_error = None
        typed = ((self.dichroic_switch and config.dichroic_mirror.type is not None) or
                 (self.emission_switch and config.emission_filter.type is not None))
        if typed:
            self.filter_error = "can't multiply sequence by non-int of type 'float'"
        self.spectral_zero = self.dichroic_switch or self.emission_switch
        self.image_magnification = config.magnification
        det = config.detector
        self.detector_type = det.type
        self.detector_image_size = tuple(int(x) for x in det.image_size)
        self.detector_pixel_length = det.pixel_length
        self.detector_exposure_time = det.exposure_time
        self.detector_focal_point = numpy.asarray(det.focal_point, dtype=float)
        self.detector_qeff = det.QE
        self.detector_readout_noise = det.readout_noise
        self.detector_emgain = det.emgain
        adc = config.analog_to_digital_converter
        self.ADConverter_bit = adc.bit
        self.ADConverter_fullwell = adc.fullwell
        self.ADConverter_fpn_type = adc.type
        self.ADConverter_fpn_count = adc.count
        self.ADConverter_offset0 = adc.offset
        self.ADConverter_offset, self.ADConverter_gain = self._adc_gain(adc.offset, rng)
        eff = config.effects
        self.background_switch = eff.background.switch
        self.background_mean = eff.background.mean
        self.quantum_yield = eff.fluorescence.quantum_yield
        self.abs_coefficient = eff.fluorescence.abs_coefficient
        self.photobleaching_switch = eff.photo_bleaching.switch
        self.photobleaching_half_life = eff.photo_bleaching.half_life
        fl = config.fluorophore
        self.radial_cutoff = fl.radial_cutoff
        self.depth_cutoff = fl.depth_cutoff

    def set_fluorophore(self, type=None, wave_length=None, normalization=None, radius=None,
                        radial_width=None, min_wave_length=None, max_wave_length=None,
                        radial_cutoff=None, depth_cutoff=None):
        """_epifm.py:822-858."""
        grid = numpy.arange(min_wave_length, max_wave_length, 1e-9, dtype=float)
        self.fluorophore_type = type
        self.fluorophore_radius = radius
        self.psf_normalization = normalization
        if type == 'Gaussian':
            ex = numpy.zeros(len(grid), dtype=float)
            em = numpy.zeros(len(grid), dtype=float)
            i_ex = i_em = numpy.abs(grid - wave_length).argmin()
            self.psf_radial_width = radial_width
        else:
            ex_rows, em_rows = io.read_fluorophore_catalog(type)
            ex = numpy.array(calculate_efficiency(ex_rows, grid))
            em = numpy.array(calculate_efficiency(em_rows, grid))
            i_ex, i_em = ex.argmax(), em.argmax()
            if wave_length is not None:
                warnings.warn('The given wave length [{}] was ignored'.format(wave_length))
            self.psf_radial_width = None
        ex[i_ex] = 100
        em[i_em] = 100
        ex /= sum(ex)   # builtin sum, as the reference
        em /= sum(em)
        self.fluoex_norm, self.fluoem_norm = ex, em
        self.psf_wavelength = grid[i_em]

    def _adc_gain(self, adc0, rng, shape=None):
        """_epifm.py:926-963 (fixed-pattern offsets drawn from rng) for a
        frame of `shape` (default: the configured image size)."""
        nw, nh = self.detector_image_size if shape is None else shape
        kind, count = self.ADConverter_fpn_type, self.ADConverter_fpn_count
        if kind == 'none':
            offset = numpy.full(nw * nh, adc0)
        elif kind == 'pixel':
            offset = numpy.rint(rng.normal(adc0, count, nw * nh))
        elif kind == 'column':
            column = rng.normal(adc0, count, nh)
            offset = numpy.rint(numpy.tile(column, (1, nw)).reshape(nh * nw))
        else:
            raise ValueError("FPN type [{}] is invalid ['pixel', 'column' or 'none']".format(kind))
        gain = (self.ADConverter_fullwell - 0.0) / (pow(2.0, self.ADConverter_bit) - offset)
        return offset.reshape([nw, nh]), gain.reshape([nw, nh])

    # ---- derived values used by the device path ----
    @property
    def pixel_length(self):
        return self.detector_pixel_length / self.image_magnification

    def snells_law(self):
        return snells_law(self.source_flux_density, self.source_wavelength, self.source_angle)

    def emit_photons(self, amplitude, unit_time):
        return get_emit_photons(amplitude, unit_time, self.abs_coefficient, self.quantum_yield,
                                self.fluorophore_radius)

    def weight_params(self, tables):
        amp, depth = self.snells_law()
        x_sec, one_minus = emit_constants(self.abs_coefficient, self.fluorophore_radius)
        p = _lib.WeightParamsT()
        p.amplitude = float(amp)
        p.penetration_depth = float(depth)
        p.x_sec = float(x_sec)
        p.quantum_yield = float(self.quantum_yield)
        p.one_minus_10_pow_minus_a = float(one_minus)
        spectral = 0.0 * self.fluoem_norm if self.spectral_zero else self.fluoem_norm
        p.spectral_psf_norm = float(numpy.sum(spectral) * self.psf_normalization)
        p.four_pi = float(4.0 * numpy.pi)
        p.focal_depth = float(self.detector_focal_point[0])
        p.depth_cutoff = float(self.depth_cutoff)
        p.res = RES
        p.cutoff_slot = tables.cutoff_index
        return p

    def detector_params(self):
        p = _lib.DetectorParamsT()
        p.qe = float(self.detector_qeff)
        p.background = float(self.background_mean) if self.background_switch else 0.0
        p.background_switch = 1 if self.background_switch else 0
        p.emgain = int(self.detector_emgain or 1)
        p.readout_noise = float(self.detector_readout_noise)
        p.fullwell = float(self.ADConverter_fullwell)
        p.adc_max = float(2 ** self.ADConverter_bit - 1)
        off, gain = self.ADConverter_offset, self.ADConverter_gain
        p.offset = float(off.flat[0])
        p.gain = float(gain.flat[0])
        return p

    def uniform_adc(self):
        """True when the ADC offset/gain maps are constant (FPN 'none')."""
        o, g = self.ADConverter_offset, self.ADConverter_gain
        return bool((o == o.flat[0]).all() and (g == g.flat[0]).all())


# --------------------------------------------------------------------------
# Device pipeline
# --------------------------------------------------------------------------

class FrameGeometry:
    """One camera frame as the C-ABI's smx_fov_t (plus derived splat sizes)."""

    def __init__(self, nw, nh, pixel_length, p0u, p0v, pix_off=0, tile_off=0, tile=None):
        lib = _lib.lib()
        ts = int(tile) if tile else lib.smx_splat_tile_size()
        f = _lib.FovT()
        f.nw, f.nh = int(nw), int(nh)
        f.tiles_r, f.tiles_c = (int(nw) + ts - 1) // ts, (int(nh) + ts - 1) // ts
        f.pix_off, f.tile_off = int(pix_off), int(tile_off)
        f.pixel_length = float(pixel_length)
        f.half_w = float(int(nw) * pixel_length * 0.5)
        f.half_h = float(int(nh) * pixel_length * 0.5)
        f.p0u, f.p0v = float(p0u), float(p0v)
        self.fov = f
        self.n_tiles = f.tiles_r * f.tiles_c


def tiles_per_emitter(pixel_length, ntab, tile=16):
    """Upper bound of the tiles one emitter's footprint can touch."""
    span = int(math.ceil(RES * (ntab - 1) / pixel_length)) + 2
    per_axis = span // tile + 2
    return per_axis * per_axis


class Splatter:
    """Weights + pixel integration for a batch of frames sharing one PSF."""

    def __init__(self, configs, tables, device):
        self.c = configs
        self.tables = tables
        self.device = device
        self.wparams = configs.weight_params(tables)
        self.status = torch.zeros(1, dtype=torch.int32, device=device)
        self.use_phase = True   # phase-table splat when the pixel is an integer number of cells

    def weights(self, depth, p_state, unit_time, budget=None, budget_slot=None, slot_map=None):
        n = depth.numel()
        w = torch.empty(n, dtype=torch.float64, device=self.device)
        key = torch.empty(n, dtype=torch.int32, device=self.device)
        ps_out = torch.empty(n, dtype=torch.float64, device=self.device)
        _lib.call('smx_emitter_weights', n, depth.data_ptr(), _device.ptr(p_state),
                  self.wparams, float(unit_time), _device.ptr(budget), _device.ptr(budget_slot),
                  _device.ptr(slot_map), 0 if slot_map is None else slot_map.numel(),
                  w.data_ptr(), key.data_ptr(), ps_out.data_ptr(), _device.stream())
        return w, key, ps_out

    def splat(self, u, v, w, key, fovs_dev, nfov, n_tiles, out, fov_id=None, accumulate=True):
        lib = _lib.lib()
        n = u.numel()
        t = self.tables
        cap = max(1, n * tiles_per_emitter(self.c.pixel_length, t.ntab))
        nbytes = lib.smx_splat_workspace_bytes(n, n_tiles, cap)
        ws = _device.ARENA.get('splat_ws', nbytes, self.device)
        K = t.K if (t.use_phase and self.use_phase) else None
        _lib.call('smx_splat_expected', n, _device.ptr(u), _device.ptr(v), _device.ptr(w),
                  _device.ptr(key), _device.ptr(fov_id), fovs_dev.data_ptr(), nfov, n_tiles,
                  t.sat.data_ptr(), t.sat_elems, len(t.slot_of), t.ntab, t.period, RES, RES * RES,
                  cap, 1 if accumulate else 0, _device.ptr(K), 0, t.m, out.data_ptr(),
                  ws.data_ptr(), ws.numel(), self.status.data_ptr(), _device.stream())


def _status_check(status):
    s = int(status.item())
    if s & 2:
        raise RuntimeError('scopyon_amd: splat tile list capacity exceeded')
    if s & 8:
        raise RuntimeError('scopyon_amd: a PSF table needed by an emitter was not built')
    if s & 4:
        raise ValueError('probabilities contain NaN')


class _EPIFMSimulator:
    """GPU counterpart of src/scopyon/_epifm.py:998-1491 (processes == 1)."""

    def __init__(self, configs, environ=None):
        self.configs = configs
        self.environ = environ
        self.device = _device.require_gpu()
        c = configs
        self.tables = PSFTables(c.fluorophore_type, c.psf_wavelength, c.radial_cutoff,
                                c.depth_cutoff, c.psf_radial_width, c.pixel_length, self.device)
        self.splatter = Splatter(c, self.tables, self.device)
        nw, nh = c.detector_image_size
        p0 = c.detector_focal_point
        self.geom = FrameGeometry(nw, nh, c.pixel_length, p0[1], p0[2])
        self.fovs_dev = torch.from_numpy(
            numpy.frombuffer(bytes(self.geom.fov), dtype=numpy.uint8).copy()).to(self.device)
        self.det = c.detector_params()
        self._maps = None if c.uniform_adc() else (
            _device.upload(c.ADConverter_gain, device=self.device),
            _device.upload(c.ADConverter_offset, device=self.device))
        # photobleaching: molecule id -> budget slot (device budgets)
        self._budget_slot = {}
        self._budget = torch.zeros(0, dtype=torch.float64, device=self.device)
        self._philox_frame = 0

    # ---- bleaching state (host dict <-> device array) ----
    def _bleach_slots(self, m_ids, states, rng):
        """Slots for m_ids; draws first-sight budgets in particle order (_epifm.py:1296-1301)."""
        new = []
        slots = numpy.empty(len(m_ids), dtype=numpy.int32)
        for i, m in enumerate(m_ids):
            s = self._budget_slot.get(m)
            if s is None:
                s = len(self._budget_slot)
                self._budget_slot[m] = s
                new.append((s, m))
            slots[i] = s
        if new:
            grow = len(self._budget_slot) - self._budget.numel()
            if grow > 0:
                self._budget = torch.cat([self._budget, torch.zeros(grow, dtype=torch.float64,
                                                                   device=self.device)])
            vals = numpy.empty(len(new))
            amp0, _ = self.configs.snells_law()
            N_emit0 = self.configs.emit_photons(amp0, 1.0)
            half_life = self.configs.photobleaching_half_life
            for j, (s, m) in enumerate(new):
                if m in states:
                    vals[j] = states[m]
                elif isinstance(rng, PhiloxRNG):
                    vals[j] = rng.host_exponential(half_life / numpy.log(2.0)) * N_emit0
                else:
                    beta = half_life / numpy.log(2.0)
                    vals[j] = rng.exponential(scale=beta, size=None) * N_emit0
            idx = _device.upload(numpy.array([s for s, _ in new]), dtype=torch.int64,
                                 device=self.device)
            self._budget[idx] = _device.upload(vals, device=self.device)
        return _device.upload(slots, dtype=torch.int32, device=self.device)

    def _states_dict(self):
        vals = self._budget.cpu().numpy()
        return {m: float(vals[s]) if vals[s] != 0 else 0 for m, s in self._budget_slot.items()}

    # ---- one sub-step ----
    def _substep(self, particles, unit_time, camera, states, rng, optinfo, full_output):
        n = len(particles)
        if n == 0:
            return
        p = numpy.ascontiguousarray(particles, dtype=numpy.float64)
        c = self.configs
        if c.filter_error is not None:
            raise TypeError(c.filter_error)
        keys = depth_keys(p[:, 0], c.detector_focal_point[0], c.depth_cutoff)
        self.tables.ensure(keys)
        dev = _device.upload(p, device=self.device)
        depth, u, v = dev[:, 0].contiguous(), dev[:, 1].contiguous(), dev[:, 2].contiguous()
        p_state = dev[:, 4].contiguous()
        budget = slots = None
        if states is not None and c.photobleaching_switch:
            m_ids = [int(m) for m in p[:, 3]]
            if len(set(m_ids)) != len(m_ids):
                raise NotImplementedError('duplicate molecule ids within one time point')
            slots = self._bleach_slots(m_ids, states, rng)
            budget = self._budget
        w, key, ps = self.splatter.weights(depth, p_state, unit_time, budget, slots,
                                           self.tables.slot_map())
        self.splatter.splat(u, v, w, key, self.fovs_dev, 1, self.geom.n_tiles, camera)
        if full_output:
            self._true_data(p, unit_time, w, ps, optinfo)

    def _true_data(self, p, unit_time, w, ps, optinfo):
        """Per-molecule sums of _epifm.py:1321-1333."""
        w = w.cpu().numpy()
        ps = ps.cpu().numpy()
        depth = p[:, 0] - self.configs.detector_focal_point[0]
        for i in range(len(p)):
            m = int(p[i, 3])
            if m not in optinfo:
                optinfo[m] = numpy.zeros(8, dtype=numpy.float64)
            optinfo[m] += numpy.array([
                unit_time, unit_time * ps[i], unit_time * p[i, 1], unit_time * p[i, 2],
                unit_time * p[i, 1], unit_time * p[i, 2], unit_time * depth[i], w[i]])

    # ---- detector ----
    def detector(self, camera, rng, full_output):
        c = self.configs
        n = camera.numel()
        if c.detector_type not in ('EMCCD', 'CCD', 'CMOS'):
            raise RuntimeError("Unknown detector type was given [{}]. ".format(c.detector_type)
                               + "Use either one of 'CMOS', 'CCD' or 'EMCCD'.")
        st = _device.stream()
        E = torch.empty(n, dtype=torch.float64, device=self.device)
        adc = torch.empty(n, dtype=torch.float64, device=self.device)
        gain_map, offset_map = (None, None) if self._maps is None else self._maps
        if c.detector_type != 'EMCCD':
            return self._photon_detector(camera, rng, E, adc, gain_map, offset_map)
        if isinstance(rng, PhiloxRNG):
            frame = self._philox_frame
            self._philox_frame += 1
            _lib.call('smx_emccd_philox', n, camera.data_ptr(), _device.ptr(gain_map),
                      _device.ptr(offset_map), self.det, rng.seed, frame, 0, None, E.data_ptr(),
                      adc.data_ptr(), 0, st)
            return E, adc
        lib = _lib.lib()
        flag = _device.ARENA.typed('det_flag', n, torch.int32, self.device)
        uidx = _device.ARENA.typed('det_uidx', n, torch.int32, self.device)
        _lib.call('smx_detector_expectation', n, camera.data_ptr(), self.det, E.data_ptr(),
                  flag.data_ptr(), st)
        sws = _device.ARENA.get('det_scan', lib.smx_scan_workspace_bytes(n), self.device)
        _lib.call('smx_exclusive_scan_i32', n, flag.data_ptr(), uidx.data_ptr(), sws.data_ptr(),
                  sws.numel(), st)
        count = int(uidx[-1].item()) + int(flag[-1].item())
        shape = c.detector_image_size
        # reference order: readout noise (one vector draw), then one uniform
        # per E > 0 pixel inside RandomState.choice (_epifm.py:1449-1454, 393-400)
        if c.detector_readout_noise <= 0:
            noise = None
        else:
            noise = _device.upload(rng.normal(0, c.detector_readout_noise, shape),
                                   device=self.device)
        uni = _device.upload(rng.random_sample(count) if count else numpy.zeros(1),
                             device=self.device)
        E_common = float(c.detector_qeff * ((0.0 + c.background_mean) if c.background_switch
                                            else 0.0))
        T_common = _device.ARENA.typed('det_tcommon', int(_lib.lib().smx_emccd_common_elems()),
                                       torch.float64, self.device)
        _lib.call('smx_emccd_common_total', E_common, self.det.emgain, T_common.data_ptr(), st)
        self.splatter.status.zero_()
        _lib.call('smx_emccd_parity', n, E.data_ptr(), _device.ptr(noise), uni.data_ptr(),
                  uidx.data_ptr(), _device.ptr(gain_map), _device.ptr(offset_map), self.det,
                  E_common, T_common.data_ptr(), adc.data_ptr(), None,
                  self.splatter.status.data_ptr(), st)
        return E, adc

    def _cmos_table(self):
        """CMOS readout-noise values and the cdf RandomState.choice builds from
        their weights (_epifm.py:331-345): p = w / sum(w); cdf = cumsum(p) / cdf[-1]."""
        if getattr(self, '_cmos', None) is None:
            table = numpy.loadtxt(os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                               'catalog', 'detector', 'RNDist_F40.csv'),
                                  delimiter=',')
            p = table[:, 1] / table[:, 1].sum()
            cdf = p.cumsum()
            cdf /= cdf[-1]
            self._cmos = (_device.upload(table[:, 0], device=self.device),
                          _device.upload(cdf, device=self.device), len(cdf))
        return self._cmos

    def _photon_detector(self, camera, rng, E, adc, gain_map, offset_map):
        """CCD and CMOS (_epifm.py:329-353, 418-433, 1443-1469): Poisson
        photo-electrons plus Gaussian (CCD) or tabulated (CMOS) readout noise."""
        c = self.configs
        n = camera.numel()
        st = _device.stream()
        cmos = c.detector_type == 'CMOS'
        tv, tc, tn = self._cmos_table() if cmos else (None, None, 0)
        if isinstance(rng, PhiloxRNG):
            frame = self._philox_frame
            self._philox_frame += 1
            _lib.call('smx_photon_detector_philox', n, camera.data_ptr(), _device.ptr(gain_map),
                      _device.ptr(offset_map), self.det, _device.ptr(tv), _device.ptr(tc), tn,
                      rng.seed, frame, 0, None, E.data_ptr(), adc.data_ptr(), 0, st)
            return E, adc
        _lib.call('smx_detector_expectation', n, camera.data_ptr(), self.det, E.data_ptr(), None,
                  st)
        shape = c.detector_image_size
        # reference order: the noise (CMOS: one RandomState.choice uniform per
        # pixel; CCD: one vector normal draw), then rng.poisson per pixel
        if cmos:
            noise = _device.upload(rng.random_sample(n), device=self.device)
        elif c.detector_readout_noise > 0:
            noise = _device.upload(rng.normal(0, c.detector_readout_noise, shape),
                                   device=self.device)
        else:
            noise = None
        signal = rng.poisson(E.cpu().numpy().reshape(shape)).astype(numpy.float64)
        sig = _device.upload(signal, device=self.device)
        _lib.call('smx_adc_parity', n, sig.data_ptr(), _device.ptr(noise), _device.ptr(tv),
                  _device.ptr(tc), tn, _device.ptr(gain_map), _device.ptr(offset_map), self.det,
                  adc.data_ptr(), st)
        return E, adc

    # ---- frames ----
    def output_frame(self, input_data, frame_index=0, start_time=0.0, exposure_time=None,
                     fluorescence_states=None, rng=None, processes=None, full_output=True):
        """_epifm.py:1121-1225."""
        c = self.configs
        exposure_time = exposure_time or c.detector_exposure_time
        if rng is None:
            rng = numpy.random.RandomState()
        times = numpy.array([t for t, _ in input_data])
        t = start_time + exposure_time * frame_index
        if c.shutter.get('switch'):
            t = max(t, c.shutter['start_time'])
            exposure_time = max(0.0, min(t + exposure_time, c.shutter['end_time']))
        start = numpy.searchsorted(times, t, side='right')
        if start != 0:
            start -= 1
        stop = numpy.searchsorted(times, t + exposure_time, side='left')
        if times[start] > t:
            warnings.warn("No data input for interval [{}, {}]".format(t, times[start]))
        frame_data = input_data[start:stop]
        nw, nh = c.detector_image_size
        camera = torch.zeros(nw * nh, dtype=torch.float64, device=self.device)
        self.splatter.status.zero_()
        optinfo = {}
        for i, (i_time, particles) in enumerate(frame_data):
            current = i_time if i != 0 else t
            nxt = frame_data[i + 1][0] if i + 1 < len(frame_data) else t + exposure_time
            unit_time = nxt - current
            if unit_time < 1e-13:
                continue
            self._substep(particles, unit_time, camera, fluorescence_states, rng, optinfo,
                          full_output)
        _status_check(self.splatter.status)
        if full_output:
            p0 = c.detector_focal_point
            pl = c.pixel_length
            for m_id in optinfo:
                optinfo[m_id][1] /= exposure_time
                optinfo[m_id][2:6] /= optinfo[m_id][0]
                optinfo[m_id][2] = (optinfo[m_id][2] - p0[1]) / pl + (nw - 1) * 0.5
                optinfo[m_id][3] = (optinfo[m_id][3] - p0[2]) / pl + (nh - 1) * 0.5
        E, adc = self.detector(camera, rng, full_output)
        _status_check(self.splatter.status)
        cam = numpy.empty((nw, nh, 2))
        cam[:, :, 0] = E.cpu().numpy().reshape(nw, nh)
        cam[:, :, 1] = adc.cpu().numpy().reshape(nw, nh)
        info = dict(true_data=optinfo)
        if fluorescence_states is not None:
            fluorescence_states.update(self._states_dict())
            info['fluorescence_states'] = copy.copy(fluorescence_states)
        return cam, info

    def generate_frames(self, input_data, num_frames, start_time=0.0, exposure_time=None,
                        rng=None, processes=None, full_output=True):
        """_epifm.py:1017-1049."""
        if rng is None:
            rng = numpy.random.RandomState()
        states = {} if self.configs.photobleaching_switch else None
        exposure_time = exposure_time or self.configs.detector_exposure_time
        for k in range(num_frames):
            yield self.output_frame(input_data, frame_index=k, start_time=start_time,
                                    exposure_time=exposure_time, fluorescence_states=states,
                                    rng=rng, processes=processes, full_output=full_output)
