"""float64 restatement of scopyon's image formation and detector path.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  Every function keeps
the reference's operation order so that, on the same numpy/scipy, it
reproduces the reference bit for bit; ``file:line`` cites ``src/scopyon``.
"""
import math
import warnings

import numpy
from scipy.interpolate import interp1d
from scipy.special import i1e, j0

from .params import get_emit_photons, snells_law

RES = 1.0e-9            # _epifm.py:59-60 (depth and radial resolution)


# --------------------------------------------------------------------------
# PSF (_epifm.py:48-213)
# --------------------------------------------------------------------------

def born_wolf_radial(r, z, wave_length):
    """_epifm.py:180-213: 100-point Riemann sum in rho, NA fixed at 1.4."""
    NA = 1.4
    k = 2.0 * numpy.pi / wave_length
    alpha = k * NA
    gamma = k * numpy.power(NA / 2, 2)
    drho = 1.0 / 100
    rho = numpy.arange(1, 100 + 1) * drho
    # j0(x * alpha * rho) for each x: (x * alpha) is formed first (:198)
    J0 = j0((r * alpha)[:, None] * rho[None, :])
    Y = numpy.exp(-2 * 1.j * z * gamma * rho * rho) * rho * drho
    I_sum = (Y * J0).sum(axis=1)
    # per-element abs()**2 on numpy scalars, as the map at :208
    psf = numpy.array([abs(v) ** 2 for v in I_sum])
    psf *= alpha * alpha / numpy.pi
    return psf


def gaussian_radial(r, radial_width):
    """_epifm.py:132-134 (no depth dependence)."""
    return numpy.exp(-0.5 * (r / radial_width) ** 2) / (2 * numpy.pi * radial_width * radial_width)


def radial_grid(radial_cutoff):
    """_epifm.py:91: radial sample points [m]."""
    return numpy.arange(0.0, radial_cutoff, RES, dtype=float)


def radial_to_cartesian(radial, psf_r):
    """_epifm.py:98-126: (2n-1)^2 table, R clamped to max(radial), linear interp."""
    n = radial.size
    off = (numpy.arange(2 * (n - 1) + 1) - (n - 1)) * RES
    R = numpy.sqrt(off[None, :] ** 2 + off[:, None] ** 2)
    R[R > radial.max()] = radial.max()
    return interp1d(radial, psf_r)(R)


def depth_key(depth, depth_cutoff):
    """_epifm.py:76-85 -> (cache key, depth used for the profile)."""
    depth = abs(depth)
    if depth < depth_cutoff + RES:
        key = int(depth / RES)
        return key, key * RES
    return -1, depth_cutoff


class PSFCache:
    """_epifm.py:48-96: per-key cartesian tables, built on first use."""

    def __init__(self, p):
        self.p = p
        self.tables = {}

    def radial(self, depth_used):
        r = radial_grid(self.p.radial_cutoff)
        if self.p.fluorophore_type == 'Gaussian':
            return r, gaussian_radial(r, self.p.psf_radial_width)
        return r, born_wolf_radial(r, depth_used, self.p.psf_wavelength)

    def get(self, depth):
        key, used = depth_key(depth, self.p.depth_cutoff)
        if key not in self.tables:
            r, psf_r = self.radial(used)
            self.tables[key] = radial_to_cartesian(r, psf_r)
        return self.tables[key]


# --------------------------------------------------------------------------
# Splat (_epifm.py:215-282)
# --------------------------------------------------------------------------

def _boundaries(n_pix, pixel_length, offset, signal_len, extent):
    """Pixel range and table-row boundaries for one axis (_epifm.py:233-253)."""
    half = n_pix * pixel_length * 0.5
    lo = math.floor((half + offset - extent * 0.5) / pixel_length)
    hi = math.ceil((half + offset + extent * 0.5) / pixel_length)
    idx = numpy.arange(max(0, lo), min(n_pix, hi) + 1)
    edges = (idx * pixel_length - (half + offset - extent * 0.5)) / RES
    edges = numpy.ceil(edges).astype(int)
    if len(edges):
        edges[0] = max(edges[0], 0)
        edges[-1] = min(edges[-1], signal_len)
    return idx, edges


def overlay_signal_(expected, signal, p_i, pixel_length, normalization):
    """_epifm.py:223-282: per pixel, sum of table cells inside the pixel."""
    _, xi, yi = p_i
    nw, nh = expected.shape
    ext_w = RES * (signal.shape[0] - 1)
    ext_h = RES * (signal.shape[1] - 1)
    iarr, left = _boundaries(nw, pixel_length, xi, signal.shape[0], ext_w)
    if len(left) == 0:
        return
    jarr, top = _boundaries(nh, pixel_length, yi, signal.shape[1], ext_h)
    if len(top) == 0:
        return
    unit_area = RES * RES
    for i, i0, i1 in zip(iarr, left[:-1], left[1:]):
        for j, j0, j1 in zip(jarr, top[:-1], top[1:]):
            photons = signal[i0:i1, j0:j1].sum() * unit_area
            if photons > 0:
                expected[i, j] += photons * normalization


def overlay_signal(psf, expected, p_i, pixel_length, normalization):
    """_epifm.py:215-221."""
    if normalization <= 0.0:
        return
    overlay_signal_(expected, psf.get(p_i[0]), p_i, pixel_length, normalization)


# --------------------------------------------------------------------------
# Detectors (_epifm.py:329-433) and ADC (_epifm.py:1472-1484)
# --------------------------------------------------------------------------

def emccd_noise(shape, readout_noise, rng):
    """_epifm.py:357-362."""
    if readout_noise <= 0:
        return numpy.zeros(shape, dtype=numpy.float64)
    return rng.normal(0, readout_noise, shape)


def emccd_probability(S, E, a):
    """_epifm.py:364-369 (scaled compound Poisson-Gamma density)."""
    X = a * S
    Y = 2 * numpy.sqrt(E * X)
    return 1.0 / Y * numpy.exp(-X + Y) * i1e(Y)


def emccd_distribution(E, emgain):
    """_epifm.py:371-391 -> (support S, normalised pmf)."""
    sigma = numpy.sqrt(E) * 5 + 10
    s_min = max(0, emgain * int(E - sigma))
    s_max = emgain * int(E + sigma)
    S = numpy.arange(s_min, s_max)
    a = 1.0 / emgain
    if S[0] > 0:
        p = emccd_probability(S, E, a)
    else:
        p = numpy.zeros(shape=(len(S),))
        p[0] = 1.0 / (2 * a * E)
        p[1:] = emccd_probability(S[1:], E, a)
    p /= p.sum()
    return S, p


def emccd_draw(E, emgain, rng, cache=None):
    """_epifm.py:393-400 (one RandomState.choice; no draw when E <= 0).
    ``cache`` (a dict) memoises the pmf per E value; the draws are unchanged."""
    if E <= 0:
        return 0
    if cache is None:
        S, p = emccd_distribution(E, emgain)
    else:
        key = float(E)
        if key not in cache:
            cache[key] = emccd_distribution(E, emgain)
        S, p = cache[key]
        p = p.copy()
    return rng.choice(S, None, p=p)


def emccd_index(E, emgain, u):
    """``RandomState.choice(S, p=p)`` restated for a given uniform ``u``:
    cdf = cumsum(p); cdf /= cdf[-1]; idx = searchsorted(cdf, u, 'right').
    Returns the drawn signal S[idx]."""
    S, p = emccd_distribution(E, emgain)
    cdf = p.cumsum()
    cdf /= cdf[-1]
    return S[cdf.searchsorted(u, side='right')]


def emccd_signal(expected, emgain, rng, memo=True):
    """_epifm.py:402-416, single-process branch (C-order pixel loop)."""
    shape = expected.shape
    flat = expected.flatten()
    out = numpy.zeros(flat.size, dtype=flat.dtype)
    cache = {} if memo else None
    for i in range(out.size):
        out[i] = emccd_draw(flat[i], emgain, rng, cache)
    return out.reshape(shape)


def emccd_signal_predrawn(expected, emgain, uniforms):
    """Same as ``emccd_signal`` with the uniforms of the E>0 pixels given in
    C order (what ``random_sample`` would return inside ``choice``)."""
    flat = expected.flatten()
    out = numpy.zeros(flat.size, dtype=flat.dtype)
    k = 0
    for i in range(out.size):
        if flat[i] > 0:
            out[i] = emccd_index(flat[i], emgain, uniforms[k])
            k += 1
    return out.reshape(expected.shape)


def ccd_noise(shape, readout_noise, rng):
    """_epifm.py:420-425."""
    if readout_noise <= 0:
        return numpy.zeros(shape, dtype=numpy.float64)
    return rng.normal(0, readout_noise, shape)


def poisson_signal(expected, rng):
    """_epifm.py:347-353 / 427-433: per-pixel rng.poisson, row-major."""
    out = numpy.zeros(expected.shape, dtype=expected.dtype)
    for i in range(expected.shape[0]):
        for j in range(expected.shape[1]):
            out[i, j] = rng.poisson(expected[i, j], None)
    return out


def cmos_noise(shape, rng, table):
    """_epifm.py:331-345 with the (value, weight) table of RNDist_F40.csv."""
    values, weights = table[:, 0], table[:, 1]
    p = weights / weights.sum()
    out = numpy.zeros(shape, dtype=numpy.float64)
    for i in range(shape[0]):
        for j in range(shape[1]):
            out[i, j] = rng.choice(values, None, p=p)
    return out


def adc_counts(photoelectron, fullwell, gain, offset, bit):
    """_epifm.py:1472-1484 (no rounding)."""
    ADC = photoelectron.copy()
    ADC[ADC > fullwell] = fullwell
    ADC_max = 2 ** bit - 1
    ADC /= gain
    ADC += offset
    ADC[ADC > ADC_max] = ADC_max
    ADC[ADC < 0] = 0
    return ADC


def detector_output(p, rng, camera, cmos_table=None):
    """_epifm.py:1430-1470 (in place on camera[:, :, 0:2])."""
    photons = camera[:, :, 0]
    if p.background_switch:
        photons += p.background_mean
    expected = p.QE * photons
    if p.detector_type == 'CMOS':
        noise = cmos_noise(expected.shape, rng, cmos_table)
        signal = poisson_signal(expected, rng)
    elif p.detector_type == 'EMCCD':
        noise = emccd_noise(expected.shape, p.readout_noise, rng)
        signal = emccd_signal(expected, p.emgain, rng)
    elif p.detector_type == 'CCD':
        noise = ccd_noise(expected.shape, p.readout_noise, rng)
        signal = poisson_signal(expected, rng)
    else:
        raise RuntimeError("Unknown detector type was given [{}]. ".format(p.detector_type))
    camera[:, :, 0] = expected
    camera[:, :, 1] = adc_counts(signal + noise, p.adc_fullwell, p.adc_gain, p.adc_offset, p.adc_bit)
    return camera


# --------------------------------------------------------------------------
# Emitters (_epifm.py:1267-1333, 1486-1491)
# --------------------------------------------------------------------------

def photon_budget(N_emit0, half_life, rng):
    """_epifm.py:1486-1491."""
    beta = half_life / numpy.log(2.0)
    return rng.exponential(scale=beta, size=None) * N_emit0


def overlay_molecule(p, psf, expected, particle, p_0, unit_time, optinfo, states, rng):
    """_epifm.py:1267-1333 for one particle row (depth, u, v, m_id, p_state)."""
    x, y, z, m_id, p_state = particle
    m_id = int(m_id)
    p_i = numpy.array([x, y, z])
    amplitude, penet_depth = snells_law(p.light_source)
    depth = (p_i - numpy.asarray(p_0))[0]
    amplitude = amplitude * numpy.exp(-abs(depth) / penet_depth)
    N_emit = get_emit_photons(amplitude, unit_time, p.abs_coefficient, p.quantum_yield,
                              p.fluorophore_radius)
    if states is not None and p.bleaching_switch:
        if m_id not in states:
            amp0, _ = snells_law(p.light_source)
            N_emit0 = get_emit_photons(amp0, 1.0, p.abs_coefficient, p.quantum_yield,
                                       p.fluorophore_radius)
            states[m_id] = photon_budget(N_emit0, p.bleaching_half_life, rng)
        budget = states[m_id] - N_emit
        if budget <= 0:
            budget = 0
            p_state = 0.0
        states[m_id] = budget
    # _epifm.py:1308-1315: fluoem_norm is scaled in place; an efficiency read
    # from a catalog is a Python list (_epifm.py:996), so 0.01 * eff raises
    # TypeError; without a type it is numpy zeros (_epifm.py:877, 980)
    normalization = p.fluoem_norm
    if p.dichroic_switch:
        normalization *= 0.01 * p.dichroic_eff
    if p.emission_switch:
        normalization *= 0.01 * p.emission_eff
    normalization = numpy.sum(normalization) * p.psf_normalization
    normalization *= p_state * N_emit / (4.0 * numpy.pi)
    pixel_length = p.pixel_length / p.magnification
    overlay_signal(psf, expected, p_i - numpy.asarray(p_0), pixel_length, normalization)
    if m_id not in optinfo:
        optinfo[m_id] = numpy.zeros(8, dtype=numpy.float64)
    optinfo[m_id] += numpy.array([
        unit_time, unit_time * p_state, unit_time * p_i[1], unit_time * p_i[2],
        unit_time * p_i[1], unit_time * p_i[2], unit_time * depth, normalization])
    return normalization


def molecule_plane(p, psf, particles, shape, p_0, unit_time, optional_info, states, rng):
    """_epifm.py:1227-1265, single-process branch."""
    expected = numpy.zeros(shape)
    optinfo = {int(m): optional_info[int(m)] for _, _, _, m, _ in particles
               if int(m) in optional_info}
    sub = None
    if states is not None:
        sub = {int(m): states[int(m)] for _, _, _, m, _ in particles if int(m) in states}
    for i in range(len(particles)):
        overlay_molecule(p, psf, expected, particles[i], p_0, unit_time, optinfo, sub, rng)
    return expected, optinfo, sub


# --------------------------------------------------------------------------
# Frames (_epifm.py:1017-1049, 1121-1225) and inputs (base.py:61-127)
# --------------------------------------------------------------------------

def format_data(inputs, pre):
    """base.py:61-110 -> (n, 5) rows (depth, u, v, m_id, p_state)."""
    if inputs.ndim != 2:
        raise ValueError("The given 'inputs' has wrong dimension.")
    n, c = inputs.shape
    scale = pre['scale']
    if c == 2:
        data = numpy.hstack((numpy.zeros((n, 1)), inputs * scale))
        data = numpy.hstack((data, numpy.zeros((n, 2))))
        data[:, 3] = numpy.arange(n)
        data[:, 4] = 1.0
    elif c in (3, 5):
        origin = numpy.array(pre['origin'])
        data = inputs[:, :3] * scale - origin
        unit_z = numpy.cross(pre['unit_x'], pre['unit_y'])
        cols = [numpy.dot(data, unit_z).reshape((-1, 1)),
                numpy.dot(data, pre['unit_x']).reshape((-1, 1)),
                numpy.dot(data, pre['unit_y']).reshape((-1, 1))]
        if c == 5:
            cols.append(inputs[:, 3:])
            data = numpy.hstack(cols)
        else:
            data = numpy.hstack(cols)
            data = numpy.hstack((data, numpy.zeros((n, 2))))
            data[:, 3] = numpy.arange(n)
            data[:, 4] = 1.0
    elif c == 4:
        data = numpy.hstack((numpy.zeros((n, 1)), inputs[:, :2] * scale, inputs[:, 2:]))
    else:
        raise ValueError("The given 'inputs' has wrong shape.")
    return data


def format_inputs(inputs, pre):
    """base.py:112-127."""
    if isinstance(inputs, numpy.ndarray):
        return ((0.0, format_data(inputs, pre)),)
    return [(t, format_data(a, pre)) for t, a in inputs]


def output_frame(p, psf, input_data, frame_index=0, start_time=0.0, exposure_time=None,
                 states=None, rng=None, cmos_table=None):
    """_epifm.py:1121-1225 (processes == 1)."""
    exposure_time = exposure_time or p.exposure_time
    times = numpy.array([t for t, _ in input_data])
    t = start_time + exposure_time * frame_index
    if p.shutter['switch']:
        t = max(t, p.shutter['start_time'])
        exposure_time = max(0.0, min(t + exposure_time, p.shutter['end_time']))
    start = numpy.searchsorted(times, t, side='right')
    if start != 0:
        start -= 1
    stop = numpy.searchsorted(times, t + exposure_time, side='left')
    if times[start] > t:
        warnings.warn("No data input for interval [{}, {}]".format(t, times[start]))
    frame_data = input_data[start:stop]
    p_0 = numpy.asarray(p.focal_point)
    nw, nh = p.image_size
    pixel_length = p.pixel_length / p.magnification
    camera = numpy.zeros((nw, nh, 2))
    optinfo = {}
    for i, (i_time, particles) in enumerate(frame_data):
        current = i_time if i != 0 else t
        nxt = frame_data[i + 1][0] if i + 1 < len(frame_data) else t + exposure_time
        unit_time = nxt - current
        if unit_time < 1e-13:
            continue
        expected, opt_, sub = molecule_plane(p, psf, particles, camera.shape[:2], p_0,
                                             unit_time, optinfo, states, rng)
        camera[:, :, 0] += expected
        optinfo.update(opt_)
        if states is not None:
            states.update(sub)
    for m_id in optinfo:
        optinfo[m_id][1] /= exposure_time
        optinfo[m_id][2:6] /= optinfo[m_id][0]
        optinfo[m_id][2] = (optinfo[m_id][2] - p_0[1]) / pixel_length + (nw - 1) * 0.5
        optinfo[m_id][3] = (optinfo[m_id][3] - p_0[2]) / pixel_length + (nh - 1) * 0.5
    camera = detector_output(p, rng, camera, cmos_table)
    info = dict(true_data=optinfo)
    if states is not None:
        info['fluorescence_states'] = dict(states)
    return camera, info


def generate_frames(p, psf, input_data, num_frames, start_time=0.0, exposure_time=None,
                    rng=None, cmos_table=None):
    """_epifm.py:1017-1049."""
    states = {} if p.bleaching_switch else None
    exposure_time = exposure_time or p.exposure_time
    for k in range(num_frames):
        yield output_frame(p, psf, input_data, frame_index=k, start_time=start_time,
                           exposure_time=exposure_time, states=states, rng=rng,
                           cmos_table=cmos_table)


def as_8bit(data, cmin=None, cmax=None, low=None, high=None):
    """image.py:99-123 (bytescale)."""
    cmin = data.min() if cmin is None else cmin
    cmax = data.max() if cmax is None else cmax
    low = 0 if low is None else low
    high = 255 if high is None else high
    cscale = cmax - cmin
    if cscale == 0.0:
        return numpy.ones(data.shape, dtype=numpy.uint8) * low
    scale = float(high - low) / cscale
    byte = (data - cmin) * scale + low
    return (byte.clip(low, high) + 0.5).astype(numpy.uint8)
