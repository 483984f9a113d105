"""EPIFMConfigs scalars (float64) from a plain, already-SI config mapping.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

``epifm_params(cfg, rng)`` restates ``EPIFMConfigs.initialize``
(``_epifm.py:762-797``) for the keys the hot path reads.  ``cfg`` is the
``default`` sub-tree of ``scopyon.yaml`` with every ``{value, units}`` leaf
already reduced to its SI magnitude (what ``Configuration.__getattr__`` returns,
``config.py:101-117``).
"""
import copy
import csv
import io
import os
import types

import numpy

N_A = 6.02214076e23                     # constants.py:12 (exact SI value)
HC = 6.62607015e-34 * 299792458.0       # constants.py:15-17

_CATALOG = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                        'scopyon_amd', 'catalog')


def _is_float(text):
    try:
        float(text)
        return True
    except ValueError:
        return False


def read_catalog(relpath, root=_CATALOG):
    """io.py:29-52: 5 header lines, then ``key`` rows followed by data rows."""
    with open(os.path.join(root, relpath)) as f:
        text = f.read()
    stream = io.StringIO(text)
    for _ in range(5):
        stream.readline()
    rows = csv.reader(stream)
    first = next(rows)
    if len(first) != 1 or _is_float(first[0]):
        raise RuntimeError('invalid catalog [{}]'.format(relpath))
    key, table = first[0], {}
    for row in rows:
        if len(row) == 1 and not _is_float(row[0]):
            key = row[0]
        elif len(row) != 0:
            table.setdefault(key, []).append(row)
    return table


def calculate_efficiency(data, wave_length):
    """_epifm.py:983-996 (integer-nm rows aligned onto the 1 nm grid)."""
    data = numpy.array(data, dtype=float)
    data = data[data[:, 0] % 1 == 0, :]
    eff = numpy.zeros(len(wave_length))
    grid_nm = numpy.round(wave_length / 1e-9).astype(int)
    on_grid = numpy.isin(grid_nm, data[:, 0])
    in_rows = numpy.isin(data[:, 0], grid_nm)
    eff[on_grid] = data[in_rows, 1]
    return eff.tolist()


def fluorophore(fl):
    """_epifm.py:822-858 -> (psf_wavelength, fluoem_norm, psf_radial_width)."""
    grid = numpy.arange(fl['min_wave_length'], fl['max_wave_length'], 1e-9, dtype=float)
    if fl['type'] == 'Gaussian':
        ex = numpy.zeros(len(grid), dtype=float)
        em = numpy.zeros(len(grid), dtype=float)
        i_ex = i_em = numpy.abs(grid - fl['wave_length']).argmin()
        width = fl['radial_width']
    else:
        table = read_catalog('fluorophore/{}.csv'.format(fl['type']))
        ex = numpy.array(calculate_efficiency(table['Excitation'], grid))
        em = numpy.array(calculate_efficiency(table['Emission'], grid))
        i_ex, i_em = ex.argmax(), em.argmax()
        width = None
    ex[i_ex] = 100
    em[i_em] = 100
    # builtin sum (sequential), as at _epifm.py:854-855
    ex /= sum(ex)
    em /= sum(em)
    return grid[i_em], em, width


def snells_law(light_source):
    """_epifm.py:1362-1428 -> (amplitude [photons/s/m^2], penetration depth [m])."""
    N_0 = light_source['flux_density'] / (HC / light_source['wave_length'])
    theta = light_source['angle']
    sin, cos = numpy.sin(theta), numpy.cos(theta)
    sin2, cos2 = sin ** 2, cos ** 2
    n_1, n_2 = 1.46, 1.384
    r = n_2 / n_1
    r2 = r ** 2
    if sin2 / r2 < 1:
        return N_0, numpy.inf
    A2_x = N_0 * (4 * cos2 * (sin2 - r2) / (r2 ** 2 * cos2 + sin2 - r2))
    A2_y = N_0 * (4 * cos2 / (1 - r2))
    A2_z = N_0 * (4 * cos2 * sin2 / (r2 ** 2 * cos2 + sin2 - r2))
    amplitude = ((A2_x + A2_z) + A2_y) / 2
    depth = light_source['wave_length'] / (4.0 * numpy.pi * numpy.sqrt(n_1 ** 2 * sin2 - n_2 ** 2))
    return amplitude, depth


def get_emit_photons(amplitude, unit_time, abs_coeff, QY, radius):
    """_epifm.py:1343-1360 (same operation order)."""
    x_sec = numpy.log(10) * abs_coeff * 0.1 / N_A
    n_abs = amplitude * x_sec * unit_time
    volume = (4.0 / 3.0) * numpy.pi * numpy.power(radius, 3)
    depth = 2.0 * radius
    A = (abs_coeff * 0.1 / N_A) * (1.0 / volume) * depth
    return QY * n_abs * (1.0 - numpy.power(10.0, -A))


def adc_gain(adc, image_size, rng):
    """_epifm.py:926-963 -> (offset map, gain map), shape image_size."""
    nw, nh = image_size
    adc0, kind, count = adc['offset'], adc['type'], adc['count']
    if kind == 'none':
        offset = numpy.full(nw * nh, adc0)
    elif kind == 'pixel':
        offset = numpy.rint(rng.normal(adc0, count, nw * nh))
    elif kind == 'column':
        column = rng.normal(adc0, count, nh)
        offset = numpy.rint(numpy.tile(column, (1, nw)).reshape(nh * nw))
    else:
        raise ValueError("FPN type [{}] is invalid ['pixel', 'column' or 'none']".format(kind))
    gain = (adc['fullwell'] - 0.0) / (pow(2.0, adc['bit']) - offset)
    return offset.reshape([nw, nh]), gain.reshape([nw, nh])


def epifm_params(cfg, rng=None):
    """Flatten the SI config sub-tree into the scalars the hot path reads."""
    cfg = copy.deepcopy(cfg)
    p = types.SimpleNamespace()
    fl = cfg['fluorophore']
    p.psf_wavelength, p.fluoem_norm, p.psf_radial_width = fluorophore(fl)
    p.fluorophore_type = fl['type']
    p.fluorophore_radius = fl['radius']
    p.psf_normalization = fl['normalization']
    p.radial_cutoff = fl['radial_cutoff']
    p.depth_cutoff = fl['depth_cutoff']
    p.shutter = cfg['shutter']
    p.light_source = cfg['light_source']
    p.dichroic_switch = cfg['dichroic_mirror']['switch']
    p.emission_switch = cfg['emission_filter']['switch']
    # _epifm.py:871-877, 974-980: a catalog type gives a list (tolist()), none
    # gives zeros; the list's values never matter (0.01 * list raises first)
    nwl = len(p.fluoem_norm)
    p.dichroic_eff = ([0.0] * nwl if cfg['dichroic_mirror'].get('type') is not None
                      else numpy.zeros(nwl))
    p.emission_eff = ([0.0] * nwl if cfg['emission_filter'].get('type') is not None
                      else numpy.zeros(nwl))
    p.magnification = cfg['magnification']
    det = cfg['detector']
    p.detector_type = det['type']
    p.image_size = tuple(det['image_size'])
    p.pixel_length = det['pixel_length']
    p.exposure_time = det['exposure_time']
    p.focal_point = det['focal_point']
    p.QE = det['QE']
    p.readout_noise = det['readout_noise']
    p.emgain = det['emgain']
    adc = cfg['analog_to_digital_converter']
    p.adc_bit = adc['bit']
    p.adc_fullwell = adc['fullwell']
    p.adc_offset, p.adc_gain = adc_gain(adc, p.image_size, rng)
    eff = cfg['effects']
    p.background_switch = eff['background']['switch']
    p.background_mean = eff['background']['mean']
    p.quantum_yield = eff['fluorescence']['quantum_yield']
    p.abs_coefficient = eff['fluorescence']['abs_coefficient']
    p.bleaching_switch = eff['photo_bleaching']['switch']
    p.bleaching_half_life = eff['photo_bleaching']['half_life']
    return p


def default_config():
    """scopyon.yaml:36-134 ``default`` sub-tree with every leaf in SI units."""
    return {
        'type': 'epifm',
        'fluorophore': {
            'type': 'Tetramethylrhodamine(TRITC)', 'radius': 20.0e-9,
            'min_wave_length': 300.0e-9, 'max_wave_length': 1000.0e-9,
            'normalization': 13564.06122650865, 'radial_cutoff': 1000.0e-9,
            'depth_cutoff': 1000.0e-9, 'radial_width': None, 'wave_length': None},
        'light_source': {
            'switch': True, 'type': 'LASER', 'wave_length': 532.0e-9,
            'flux_density': 400000.0, 'angle': 1.2566370614359172, 'radius': 20.0e-6},
        'dichroic_mirror': {'switch': False, 'type': 'FF562-Di03-25x36'},
        'magnification': 241.0,
        'detector': {
            'switch': True, 'type': 'EMCCD', 'image_size': [512, 512],
            'pixel_length': 16.0e-6, 'exposure_time': 0.100,
            'focal_point': [0.0, 0.0, 0.0], 'QE': 0.92, 'dark_count': 0,
            'readout_noise': 100, 'emgain': 300},
        'analog_to_digital_converter': {
            'bit': 16, 'offset': 2000, 'type': 'none', 'count': 0.0, 'fullwell': 800000},
        'excitation_filter': {'switch': False, 'type': None},
        'emission_filter': {'switch': False, 'type': None},
        'shutter': {'switch': False, 'start_time': 0.0, 'end_time': None},
        'effects': {
            'background': {'switch': True, 'mean': 0.01},
            'fluorescence': {'quantum_yield': 0.61, 'abs_coefficient': 83400},
            'photo_bleaching': {'switch': True, 'half_life': 2.5}},
    }
