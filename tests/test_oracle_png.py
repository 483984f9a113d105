"""The oracle's whole frame path against the reference's own example output.

``tests/golden/tirf_000_u8.npz`` holds the pixels of the reference's
``examples/tirf_000.png``, written by ``examples/tirf.py``: RandomState(123),
100 emitters uniform over the 512 x 512 frame, default configuration, 33 ms
exposure, ``Image.as_8bit`` (image.py:99-123).

What this pins:
* background-only pixels: the random-draw order of one frame (emitter
  uniforms, then the readout-noise vector, then one ``choice`` uniform per
  pixel, _epifm.py:355-416, 1430-1470) and the background -> QE -> EMCCD ->
  ADC chain, pixel for pixel;
* pixels under emitters: the emitter brightness.  Given the reference's own
  noise and uniforms, each bright pixel's grey level fixes the expectation
  that produced it to about 1.4% (one grey level), and the median over 400
  pixels to ~0.15%: the oracle's weight chain agrees with the image to
  better than 0.5%, and a weight off by 1% does not.

The 8-bit rescale's upper bound (the brightest pixel) depends on emitter
brightness, so the slope of the linear rescale is fitted on the background
pixels and the offset is the frame minimum, as in ``as_8bit``.
"""
import os

import numpy
import pytest

from oracle import epifm as O
from oracle import params as P

GOLDEN = os.path.join(os.path.dirname(__file__), 'golden')
pytestmark = pytest.mark.slow


def _tirf_frame():
    cfg = P.default_config()
    cfg['detector']['exposure_time'] = 33.0e-3
    rng = numpy.random.RandomState(123)
    pl = cfg['detector']['pixel_length'] / cfg['magnification']
    half = cfg['detector']['image_size'][0] * pl * 0.5
    inputs = rng.uniform(-half, +half, size=(100, 2))
    p = P.epifm_params(cfg, rng)
    psf = O.PSFCache(p)
    pre = {'scale': 1.0, 'origin': [0.0, 0, 0], 'unit_x': [1.0, 0, 0], 'unit_y': [0, 1.0, 0]}
    data = O.format_inputs(inputs, pre)
    cam, _ = O.output_frame(p, psf, list(data), rng=rng)
    return p, cam


def _background_agreement(E, adc, png):
    """Fraction of background-only pixels whose 8-bit value equals the PNG's,
    and the largest difference in grey levels."""
    bg = E == E.min()
    x = adc[bg] - adc.min()
    y = png[bg].astype(float)
    slope = (x * y).sum() / (x * x).sum()
    pred = numpy.floor(numpy.clip(x * slope, 0, 255) + 0.5)
    return (pred == y).mean(), numpy.abs(pred - y).max(), bg.sum()


@pytest.fixture(scope='module')
def frame(monkeypatch_module):
    seen = {}
    real = O.detector_output

    def spy(p, rng, camera, cmos_table=None):
        seen['camera'] = camera.copy()
        seen['state'] = rng.get_state()
        return real(p, rng, camera, cmos_table)

    monkeypatch_module.setattr(O, 'detector_output', spy)
    p, cam = _tirf_frame()
    monkeypatch_module.setattr(O, 'detector_output', real)
    return p, cam, seen


@pytest.fixture(scope='module')
def monkeypatch_module():
    mp = pytest.MonkeyPatch()
    yield mp
    mp.undo()


@pytest.fixture(scope='module')
def png():
    return numpy.load(os.path.join(GOLDEN, 'tirf_000_u8.npz'))['image']


def test_tirf_png_background_pixels(frame, png):
    _, cam, _ = frame
    exact, worst, n_bg = _background_agreement(cam[:, :, 0], cam[:, :, 1], png)
    assert n_bg > 150_000
    assert exact > 0.99, exact
    assert worst <= 1


def test_tirf_png_catches_swapped_draw_order(frame, png):
    """A detector that draws the EMCCD uniforms before the readout noise
    (the reverse of _epifm.py:1457-1458) no longer matches the PNG."""
    p, cam, seen = frame
    rng = numpy.random.RandomState()
    rng.set_state(seen['state'])
    camera = seen['camera'].copy()
    photons = camera[:, :, 0] + p.background_mean
    expected = p.QE * photons
    signal = O.emccd_signal(expected, p.emgain, rng)
    noise = O.emccd_noise(expected.shape, p.readout_noise, rng)
    adc = O.adc_counts(signal + noise, p.adc_fullwell, p.adc_gain, p.adc_offset, p.adc_bit)
    exact, worst, _ = _background_agreement(expected, adc, png)
    assert exact < 0.5, exact


def _twocolor_frames():
    """examples/twocolor.py: RandomState(123), 250 emitters, default
    configuration (100 ms exposure); form_image on emitters [0, 200) then on
    [150, 250) with the same RandomState, as the red and green channels."""
    cfg = P.default_config()
    rng = numpy.random.RandomState(123)
    pl = cfg['detector']['pixel_length'] / cfg['magnification']
    half = cfg['detector']['image_size'][0] * pl * 0.5
    inputs = rng.uniform(-half, +half, size=(250, 2))
    pre = {'scale': 1.0, 'origin': [0.0, 0, 0], 'unit_x': [1.0, 0, 0], 'unit_y': [0, 1.0, 0]}
    frames, psf = [], None
    for part in (inputs[:200], inputs[150:]):
        p = P.epifm_params(cfg, rng)          # a new simulator per form_image call
        psf = psf or O.PSFCache(p)            # same optics: the tables are reused
        cam, _ = O.output_frame(p, psf, list(O.format_inputs(part, pre)), rng=rng)
        frames.append(cam)
    return frames


def test_twocolor_png_background_pixels():
    """Both channels of the reference's examples/twocolor_000.png: the second
    frame only matches if the first consumed exactly the reference's number of
    draws (noise vector, one choice uniform per pixel) from the shared
    RandomState.  Blue is empty."""
    png = numpy.load(os.path.join(GOLDEN, 'twocolor_000_u8.npz'))['image']
    assert not png[:, :, 2].any()
    for ch, cam in enumerate(_twocolor_frames()):
        exact, worst, n_bg = _background_agreement(cam[:, :, 0], cam[:, :, 1], png[:, :, ch])
        assert n_bg > 100_000, (ch, n_bg)
        assert exact > 0.99, (ch, exact)
        assert worst <= 1, (ch, worst)


def _emitter_scale(p, seen, adc, png, E):
    """Per-pixel ratio E_png / E for the 400 brightest pixels under emitters.

    The readout noise and the per-pixel ``choice`` uniforms are the
    reference's draws (same RandomState stream).  The 8-bit slope is fitted
    on the background pixels, which do not depend on emitter brightness; a
    pixel's grey level then gives its ADC value to half a level, hence its
    EM-register output S = (ADC - offset) * gain - noise, and E_png is the
    expectation at which the EMCCD quantile function at that pixel's uniform
    returns S (linear between the quantiles at 0.97 E, E and 1.03 E)."""
    rng = numpy.random.RandomState()
    rng.set_state(seen['state'])
    noise = O.emccd_noise(E.shape, p.readout_noise, rng)
    u = rng.random_sample(E.size).reshape(E.shape)     # every pixel has E > 0
    E_bg = p.QE * (seen['camera'][:, :, 0] + p.background_mean)
    bg = E_bg == E_bg.min()
    x = adc[bg] - adc.min()
    slope = (x * png[bg]).sum() / (x * x).sum()
    cand = numpy.argwhere((E_bg > 5.0) & (png < 255))
    top = cand[numpy.argsort(-E_bg[tuple(cand.T)], kind='stable')[:400]]
    gain = numpy.broadcast_to(numpy.asarray(p.adc_gain, dtype=float).reshape(-1), (E.size,))
    offset = numpy.broadcast_to(numpy.asarray(p.adc_offset, dtype=float).reshape(-1), (E.size,))
    gain, offset = gain.reshape(E.shape), offset.reshape(E.shape)
    ratios = []
    d = 0.03
    for i, j in top:
        S_png = ((png[i, j] / slope + adc.min()) - offset[i, j]) * gain[i, j] - noise[i, j]
        e = E[i, j]
        q = [O.emccd_index(e * (1 + k * d), p.emgain, u[i, j]) for k in (-1, 0, 1)]
        if S_png >= q[1]:
            f = 1 + d * (S_png - q[1]) / max(q[2] - q[1], 1.0)
        else:
            f = 1 - d * (q[1] - S_png) / max(q[1] - q[0], 1.0)
        ratios.append(f)
    return numpy.array(ratios)


def test_tirf_png_emitter_brightness(frame, png):
    """Pixels under emitters: the oracle's emitter weight chain (snells_law x
    get_emit_photons x spectral sum x psf_normalization / 4 pi,
    _epifm.py:1280-1315, 1343-1428) reproduces the brightness the reference
    rendered into examples/tirf_000.png: the median of E_png / E over the 400
    brightest emitter pixels is 1 to within 0.5%; a weight 1% too large is
    outside that band."""
    p, cam, seen = frame
    E = p.QE * (seen['camera'][:, :, 0] + p.background_mean)
    r = _emitter_scale(p, seen, cam[:, :, 1], png, E)
    assert len(r) == 400
    assert abs(numpy.median(r) - 1.0) < 0.005, numpy.median(r)
    planted = p.QE * (1.01 * seen['camera'][:, :, 0] + p.background_mean)
    rp = _emitter_scale(p, seen, cam[:, :, 1], png, planted)
    assert abs(numpy.median(rp) - 1.0) > 0.005, numpy.median(rp)
