"""GPU weights + splat vs the oracle's overlay loop (same positions)."""
import numpy
import pytest
import torch

from oracle import epifm as O
from conftest import oracle_params, product_config

pytestmark = pytest.mark.gpu


def _oracle_expected(p, particles, unit_time):
    psf = O.PSFCache(p)
    nw, nh = p.image_size
    exp, _, _ = O.molecule_plane(p, psf, particles, (nw, nh), numpy.asarray(p.focal_point),
                                 unit_time, {}, None, None)
    return exp


def _gpu_expected(gpu, overrides, particles, unit_time):
    from scopyon_amd import _epifm
    cfg = product_config(overrides)
    sim = _epifm._EPIFMSimulator(_epifm.EPIFMConfigs(cfg.default, rng=numpy.random.RandomState(0)))
    nw, nh = sim.configs.detector_image_size
    camera = torch.zeros(nw * nh, dtype=torch.float64, device=gpu)
    sim._substep(particles, unit_time, camera, None, None, {}, False)
    _epifm._status_check(sim.splatter.status)
    return camera.cpu().numpy().reshape(nw, nh)


def _particles(rng, n, half, depth=None):
    xy = rng.uniform(-half, half, size=(n, 2))
    d = numpy.zeros(n) if depth is None else depth
    return numpy.column_stack([d, xy, numpy.arange(n), numpy.ones(n)])


CASES = [
    # (overrides, n, half-extent factor, depth spread [m])
    ({'detector': {'image_size': [64, 64]}, 'magnification': 100.0}, 40, 0.6, None),
    ({'detector': {'image_size': [48, 80]}, 'magnification': 241.0}, 30, 0.7, None),
    ({'detector': {'image_size': [40, 40]}, 'magnification': 100.0}, 25, 1.2, None),  # edges/outside
    ({'detector': {'image_size': [32, 48]}, 'magnification': 241.0}, 12, 0.6, 1.5e-6),  # 3D keys
    ({'detector': {'image_size': [33, 17]}, 'magnification': 60.0}, 10, 0.8, 2e-7),
]


@pytest.mark.parametrize('case', range(len(CASES)))
def test_splat_matches_oracle(gpu, case):
    ov, n, f, dz = CASES[case]
    rng = numpy.random.RandomState(100 + case)
    p = oracle_params(ov)
    pl = p.pixel_length / p.magnification
    half = f * 0.5 * max(p.image_size) * pl
    depth = None if dz is None else rng.uniform(-dz, dz, n)
    parts = _particles(rng, n, half, depth)
    ref = _oracle_expected(p, parts, 0.033)
    got = _gpu_expected(gpu, ov, parts, 0.033)
    scale = ref.max()
    assert scale > 0
    numpy.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-13 * scale)


def test_splat_gaussian_psf(gpu):
    ov = {'detector': {'image_size': [40, 40]}, 'magnification': 100.0,
          'fluorophore': {'type': 'Gaussian', 'radial_width': 1.5e-7, 'wave_length': 6.0e-7}}
    rng = numpy.random.RandomState(5)
    p = oracle_params(ov)
    parts = _particles(rng, 15, 0.5 * 40 * 1.6e-7)
    ref = _oracle_expected(p, parts, 0.05)
    got = _gpu_expected(gpu, ov, parts, 0.05)
    numpy.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-13 * ref.max())


def test_splat_empty_and_zero_weight(gpu):
    ov = {'detector': {'image_size': [32, 32]}, 'magnification': 100.0}
    parts = numpy.array([[0.0, 0.0, 0.0, 0, 0.0],        # p_state 0 -> weight 0, skipped
                         [0.0, 1e-3, 1e-3, 1, 1.0]])     # far outside the frame
    got = _gpu_expected(gpu, ov, parts, 0.033)
    assert not got.any()
    got = _gpu_expected(gpu, ov, numpy.zeros((0, 5)), 0.033)
    assert not got.any()


def test_splat_deterministic(gpu):
    ov = {'detector': {'image_size': [64, 64]}, 'magnification': 100.0}
    rng = numpy.random.RandomState(9)
    parts = _particles(rng, 400, 0.5 * 64 * 1.6e-7)
    a = _gpu_expected(gpu, ov, parts, 0.033)
    b = _gpu_expected(gpu, ov, parts, 0.033)
    assert numpy.array_equal(a, b)
