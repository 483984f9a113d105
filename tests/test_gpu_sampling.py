"""GPU samplers: bit-exact against the reference's own outputs (golden
vectors) and the oracle; Philox mode checked statistically."""
import os

import numpy
import pytest

from oracle import sampling as S

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), 'golden', 'sampling_golden.npz')


def test_sample_inputs_golden(gpu):
    import scopyon_amd
    g = numpy.load(GOLD)
    t = numpy.arange(0, 6) * 0.033
    r = scopyon_amd.sample_inputs(t, N=30, lower=-15e-6, upper=15e-6, ndim=2, D=0.1e-12,
                                  rng=numpy.random.RandomState(123))
    assert numpy.array_equal(numpy.stack([b for _, b in r]), g['inputs2d_x'])
    r = scopyon_amd.sample_inputs([0.02, 0.0, 0.01, 0.01, 0.05], N=12, lower=[0, 0, -1e-6],
                                  upper=[5e-6, 4e-6, 1e-6], ndim=3, D=[1e-12, 2e-12, 0.5e-12],
                                  start=7, rng=numpy.random.RandomState(7))
    assert numpy.array_equal(numpy.stack([b for _, b in r]), g['inputs3d_x'])
    assert numpy.array_equal(numpy.array([a for a, _ in r]), g['inputs3d_t'])


def test_sample_golden(gpu):
    import scopyon_amd
    g = numpy.load(GOLD)
    r = scopyon_amd.sample(numpy.array([0.0, 0.1, 0.2, 0.2, 0.35]), N=[5, 3, 2],
                           lower=[0.0, -1.0], upper=[1.0, 1.0], D=[4e-3, 1e-3, 0.0], ndim=2,
                           periodic=True, rng=numpy.random.RandomState(5))
    assert numpy.array_equal(numpy.stack(r), g['sample_periodic_x'])
    r = scopyon_amd.sample([0.0, 0.1, 0.2, 0.3], N=[5, 3, 2], D=[4e-3, 1e-3, 0.0],
                           transmat=[[0.0, 1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]], ndim=2,
                           rng=numpy.random.RandomState(11))
    assert numpy.array_equal(numpy.stack(r), g['sample_trans_x'])
    r = scopyon_amd.sample([0.0, 0.5, 1.0], N=8, lower=[1.0, 1.0, 1.0], upper=[0.0, 0.0, 0.5],
                           D=0.01, ndim=3, rng=numpy.random.RandomState(3))
    assert numpy.array_equal(numpy.stack(r), g['sample_3d_x'])


def test_sample_larger_matches_oracle(gpu):
    import scopyon_amd
    t = numpy.linspace(0, 1, 21)
    a = scopyon_amd.sample(t, N=[300, 200], lower=-1, upper=1, D=[0.01, 0.002], ndim=3,
                           periodic=True, rng=numpy.random.RandomState(17))
    b = S.sample(t, N=[300, 200], lower=-1, upper=1, D=[0.01, 0.002], ndim=3, periodic=True,
                 rng=numpy.random.RandomState(17))
    assert numpy.array_equal(numpy.stack(a), numpy.stack(b))


def test_reflect_keeps_particles_inside(gpu):
    import scopyon_amd
    t = numpy.linspace(0, 10, 51)
    r = scopyon_amd.sample(t, N=2000, lower=0.0, upper=1.0, D=0.05, ndim=2, boundary='reflect',
                           rng=numpy.random.RandomState(4))
    x = numpy.stack(r)[:, :, :2]
    assert (x >= 0).all() and (x <= 1).all()
    # long-time distribution is uniform: mean 1/2, variance 1/12
    last = x[-1].ravel()
    assert abs(last.mean() - 0.5) < 0.02 and abs(last.var() - 1 / 12) < 0.01


def test_philox_msd_slope(gpu):
    import scopyon_amd
    D, dt, nsteps, n = 0.1e-12, 0.033, 20, 20000
    t = numpy.arange(nsteps + 1) * dt
    r = scopyon_amd.sample(t, N=n, lower=-1e-3, upper=1e-3, D=D, ndim=3,
                           rng=scopyon_amd.PhiloxRNG(99))
    x = numpy.stack(r)[:, :, :3]
    msd = ((x - x[0]) ** 2).sum(axis=2).mean(axis=1)
    slope = numpy.polyfit(t, msd, 1)[0]
    assert abs(slope / (2 * 3 * D) - 1) < 0.03
    # increments are N(0, 2 D dt) per axis (KS)
    from scipy import stats
    inc = (x[1] - x[0]).ravel() / numpy.sqrt(2 * D * dt)
    assert stats.kstest(inc, 'norm').pvalue > 1e-3
