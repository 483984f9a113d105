"""Parity at the headline configuration: 64 fields of view of 512 x 512
pixels (16 um / 100x = 160 nm) with 10,000 emitters each, the folded
float32 phase table and autotuned tiles, exactly as the benchmark runs them;
and the drop-in form_image on the plain reference configuration (512 x 512,
1,000 emitters, RandomState draws)."""
import numpy
import pytest

from oracle import epifm as O
from conftest import oracle_params, product_config, PRE

pytestmark = pytest.mark.gpu

PL = 16e-6 / 100.0


def _cfg(size, exposure=0.033):
    return product_config({'magnification': 100.0,
                           'detector': {'image_size': [size, size], 'exposure_time': exposure},
                           'effects': {'photo_bleaching': {'switch': False}}})


def test_headline_batch_crop_matches_oracle(gpu):
    """Noise-free expectation E = QE (photons + background) of a 128 x 128
    crop of two FOVs of the 64-FOV batch, against the float64 oracle fed with
    every emitter of the FOV whose footprint reaches the crop: max pointwise
    relative error <= 1e-5 (float32 table, float64 sums)."""
    from scopyon_amd.batched import BatchedSimulator
    print('headline parity: 64 x 512^2, 10000 emitters per FOV', flush=True)
    S, F, N = 512, 64, 10_000
    sim = BatchedSimulator(_cfg(S), [(S, S)] * F, N, D=0.1e-12, dt=0.033, seed=1234,
                           table_precision='folded', keep_photons=True)
    sim.step()                        # one full step (move, weights, fused splat + detector)
    sim.form()                        # and the photon image of the moved emitters
    assert sim.status() & ~1 == 0
    u, v = sim.u.cpu().numpy(), sim.v.cpu().numpy()
    p = oracle_params({'detector': {'image_size': [S, S]}, 'magnification': 100.0})
    cache = O.PSFCache(p)
    r0, c0, C = 200, 300, 128
    half = S * PL * 0.5
    worst = 0.0
    for f in (0, 37):
        a, b = sim.emitter_offsets[f], sim.emitter_offsets[f + 1]
        iu = (u[a:b] + half) / PL
        iv = (v[a:b] + half) / PL
        near = (iu > r0 - 8) & (iu < r0 + C + 8) & (iv > c0 - 8) & (iv < c0 + C + 8)
        k = int(near.sum())
        assert k > 500
        parts = numpy.column_stack([numpy.zeros(k), u[a:b][near], v[a:b][near], numpy.arange(k),
                                    numpy.ones(k)])
        ref, _, _ = O.molecule_plane(p, cache, parts, (S, S), numpy.zeros(3), 0.033, {}, None,
                                     None)
        got = sim.frame(f, 'photons').cpu().numpy()
        E_ref = p.QE * (ref[r0:r0 + C, c0:c0 + C] + p.background_mean)
        E_got = p.QE * (got[r0:r0 + C, c0:c0 + C] + p.background_mean)
        rel = numpy.abs(E_got - E_ref) / E_ref
        worst = max(worst, rel.max())
    print('headline crop: max relative error {:.2e}'.format(worst), flush=True)
    assert worst <= 1e-5


def test_plain_config_form_image_512(gpu):
    """The drop-in form_image on the plain reference geometry (512 x 512,
    160 nm pixels, 1,000 emitters, TRITC, EMCCD gain 300, background on):
    expectation within 1e-10 of the oracle on every pixel, and ADC counts
    bit-exact on 4,000 pixels (the brightest 1,000 and 3,000 at random),
    each redrawn by the oracle from the same RandomState stream (readout
    noise vector, then one choice uniform per pixel)."""
    import scopyon_amd
    print('plain config: form_image 512^2, 1000 emitters', flush=True)
    S, N = 512, 1000
    rng = numpy.random.RandomState(2024)
    inputs = rng.uniform(-0.5 * S * PL, 0.5 * S * PL, size=(N, 2))
    img, info = scopyon_amd.form_image(inputs, config=_cfg(S),
                                       rng=numpy.random.RandomState(77), full_output=True)
    ov = {'detector': {'image_size': [S, S], 'exposure_time': 0.033}, 'magnification': 100.0,
          'effects': {'photo_bleaching': {'switch': False}}}
    orng = numpy.random.RandomState(77)
    p = oracle_params(ov, orng)
    data = list(O.format_inputs(inputs, PRE))
    exp, _, _ = O.molecule_plane(p, O.PSFCache(p), data[0][1], (S, S), numpy.zeros(3), 0.033, {},
                                 None, None)
    E = p.QE * (exp + p.background_mean)
    numpy.testing.assert_allclose(info['expectation'], E, rtol=1e-10, atol=0)
    noise = O.emccd_noise(E.shape, p.readout_noise, orng)
    u = orng.random_sample(E.size).reshape(E.shape)
    flat = numpy.argsort(-E, axis=None)
    pick = numpy.concatenate([flat[:1000], numpy.random.RandomState(1).choice(flat[1000:], 3000,
                                                                              replace=False)])
    adc = img.as_array()
    gain = numpy.broadcast_to(numpy.asarray(p.adc_gain, dtype=float).reshape(-1), (E.size,))
    offset = numpy.broadcast_to(numpy.asarray(p.adc_offset, dtype=float).reshape(-1), (E.size,))
    for idx in pick:
        i, j = divmod(int(idx), S)
        sig = O.emccd_index(E[i, j], p.emgain, u[i, j])
        want = O.adc_counts(numpy.array([sig + noise[i, j]]), p.adc_fullwell, gain[idx],
                            offset[idx], p.adc_bit)[0]
        assert adc[i, j] == want, (i, j, adc[i, j], want)
