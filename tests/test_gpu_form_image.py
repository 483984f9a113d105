"""End-to-end parity: scopyon_amd.form_image / generate_images vs the oracle
on the same inputs and the same RandomState stream."""
import numpy
import pytest

from oracle import epifm as O
from conftest import oracle_params, product_config, PRE

pytestmark = pytest.mark.gpu


def _oracle_frames(overrides, inputs, seed, nframes=None, exposure=None):
    rng = numpy.random.RandomState(seed)
    p = oracle_params(overrides, rng)
    data = list(O.format_inputs(inputs, PRE))
    psf = O.PSFCache(p)
    if nframes is None:
        return [O.output_frame(p, psf, data, rng=rng, exposure_time=exposure)]
    return list(O.generate_frames(p, psf, data, nframes, rng=rng, exposure_time=exposure))


@pytest.mark.parametrize('mag,shape,n', [(100.0, (48, 48), 30), (241.0, (40, 56), 25)])
def test_form_image_matches_oracle(gpu, mag, shape, n):
    import scopyon_amd
    ov = {'detector': {'image_size': list(shape), 'exposure_time': 0.033}, 'magnification': mag}
    pl = 16e-6 / mag
    rng = numpy.random.RandomState(42)
    inputs = rng.uniform(-0.5 * shape[0] * pl, 0.5 * shape[0] * pl, size=(n, 2))
    img, info = scopyon_amd.form_image(inputs, config=product_config(ov),
                                       rng=numpy.random.RandomState(5), full_output=True)
    (cam, oinfo), = _oracle_frames(ov, inputs, 5)
    numpy.testing.assert_allclose(info['expectation'], cam[:, :, 0], rtol=1e-10, atol=0)
    assert numpy.array_equal(img.as_array(), cam[:, :, 1])
    assert sorted(info['true_data']) == sorted(oinfo['true_data'])
    for m in oinfo['true_data']:
        numpy.testing.assert_allclose(info['true_data'][m], oinfo['true_data'][m], rtol=1e-12)


def test_form_image_3d_inputs(gpu):
    import scopyon_amd
    ov = {'detector': {'image_size': [32, 32], 'exposure_time': 0.05}, 'magnification': 241.0,
          'light_source': {'angle': 0.0}}
    rng = numpy.random.RandomState(3)
    pl = 16e-6 / 241.0
    inputs = numpy.column_stack([rng.uniform(-16 * pl, 16 * pl, (12, 2)),
                                 rng.uniform(-6e-7, 6e-7, 12)])
    # 3-column input -> (depth = z, u = x, v = y) through the unit vectors
    img, info = scopyon_amd.form_image(inputs, config=product_config(ov),
                                       rng=numpy.random.RandomState(1), full_output=True)
    (cam, _), = _oracle_frames(ov, inputs, 1)
    numpy.testing.assert_allclose(info['expectation'], cam[:, :, 0], rtol=1e-10)
    assert numpy.array_equal(img.as_array(), cam[:, :, 1])


def test_generate_images_bleaching_matches_oracle(gpu):
    import scopyon_amd
    from oracle import sampling as S
    ov = {'detector': {'image_size': [32, 32], 'exposure_time': 0.033}, 'magnification': 100.0,
          'effects': {'photo_bleaching': {'switch': True, 'half_life': 0.05}}}
    pl = 16e-6 / 100.0
    dt = 0.033
    t = numpy.arange(0, 7) * dt / 2          # two sub-steps per frame
    inputs = S.sample_inputs(t, N=20, lower=-14 * pl, upper=14 * pl, ndim=2, D=0.1e-12,
                             rng=numpy.random.RandomState(2))
    frames = list(scopyon_amd.generate_images(inputs, 3, config=product_config(ov),
                                              rng=numpy.random.RandomState(9), full_output=True))
    ref = _oracle_frames(ov, inputs, 9, nframes=3)
    for (img, info), (cam, oinfo) in zip(frames, ref):
        numpy.testing.assert_allclose(info['expectation'], cam[:, :, 0], rtol=1e-10)
        assert numpy.array_equal(img.as_array(), cam[:, :, 1])
        st, ost = info['fluorescence_states'], oinfo['fluorescence_states']
        assert sorted(st) == sorted(ost)
        for m in ost:
            assert st[m] == pytest.approx(ost[m], rel=1e-12, abs=0)


def test_form_image_errors(gpu):
    import scopyon_amd
    cfg = product_config({'detector': {'image_size': [8, 8]}})
    with pytest.raises(ValueError):
        scopyon_amd.form_image(numpy.zeros((3, 6)), config=cfg, rng=numpy.random.RandomState(0))
    with pytest.raises(ValueError):
        scopyon_amd.form_image(numpy.zeros(3), config=cfg, rng=numpy.random.RandomState(0))
    with pytest.raises(TypeError):
        scopyon_amd.form_image(3.0, config=cfg, rng=numpy.random.RandomState(0))
