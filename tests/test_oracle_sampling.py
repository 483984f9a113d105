"""The sampling oracle against the reference's own outputs (golden vectors
written by tests/golden/make_sampling_golden.py from src/scopyon/sampling*.py)."""
import os
import warnings

import numpy

from oracle import sampling as S

GOLD = os.path.join(os.path.dirname(__file__), 'golden', 'sampling_golden.npz')


def test_sample_inputs_golden():
    g = numpy.load(GOLD)
    t = numpy.arange(0, 6) * 0.033
    r = S.sample_inputs(t, N=30, lower=-15e-6, upper=15e-6, ndim=2, D=0.1e-12,
                        rng=numpy.random.RandomState(123))
    assert numpy.array_equal(numpy.stack([b for _, b in r]), g['inputs2d_x'])
    assert numpy.array_equal(numpy.array([a for a, _ in r]), g['inputs2d_t'])
    r = S.sample_inputs([0.02, 0.0, 0.01, 0.01, 0.05], N=12, lower=[0, 0, -1e-6],
                        upper=[5e-6, 4e-6, 1e-6], ndim=3, D=[1e-12, 2e-12, 0.5e-12], start=7,
                        rng=numpy.random.RandomState(7))
    assert numpy.array_equal(numpy.stack([b for _, b in r]), g['inputs3d_x'])
    assert numpy.array_equal(numpy.array([a for a, _ in r]), g['inputs3d_t'])


def test_sample_golden():
    g = numpy.load(GOLD)
    r = S.sample(numpy.array([0.0, 0.1, 0.2, 0.2, 0.35]), N=[5, 3, 2], lower=[0.0, -1.0],
                 upper=[1.0, 1.0], D=[4e-3, 1e-3, 0.0], ndim=2, periodic=True,
                 rng=numpy.random.RandomState(5))
    assert numpy.array_equal(numpy.stack(r), g['sample_periodic_x'])
    with warnings.catch_warnings():
        warnings.simplefilter('ignore')
        r = S.sample([0.0, 0.1, 0.2, 0.3], N=[5, 3, 2], D=[4e-3, 1e-3, 0.0],
                     transmat=[[0.0, 1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]], ndim=2,
                     rng=numpy.random.RandomState(11))
    assert numpy.array_equal(numpy.stack(r), g['sample_trans_x'])
    r = S.sample([0.0, 0.5, 1.0], N=8, lower=[1.0, 1.0, 1.0], upper=[0.0, 0.0, 0.5], D=0.01,
                 ndim=3, rng=numpy.random.RandomState(3))
    assert numpy.array_equal(numpy.stack(r), g['sample_3d_x'])


def test_planted_mistake_is_caught():
    """A dim-major instead of particle-major draw order must fail the golden check."""
    g = numpy.load(GOLD)
    rng = numpy.random.RandomState(123)
    t = numpy.arange(0, 6) * 0.033
    pts, _ = S.sample_points(rng, N=30, lower=-15e-6, upper=15e-6, ndim=2)
    xs = [pts.copy()]
    for _ in range(5):
        z = rng.normal(0.0, numpy.sqrt(2 * 0.1e-12 * 0.033), size=(2, 30)).T  # wrong order
        pts = pts.copy()
        pts[:, :2] += z
        xs.append(pts)
    assert not numpy.array_equal(numpy.stack(xs), g['inputs2d_x'])


def test_normal_stream_identity():
    """rng.normal(0, s) per scalar == 0 + s * standard_normal vector (same gauss stream)."""
    a = numpy.random.RandomState(7)
    v = numpy.array([a.normal(0.0, 2.5) for _ in range(9)])
    b = numpy.random.RandomState(7)
    w = 0.0 + 2.5 * b.standard_normal(9)
    assert numpy.array_equal(v, w)
