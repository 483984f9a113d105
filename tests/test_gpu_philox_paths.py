"""Device-randomness (PhiloxRNG) forms of the paths the reference runs on a
RandomState: state transitions and per-state D (sampling2.py:52-68,
114-120), per-axis D (sampling.py:115-118), uniform initial positions
(sampling2.py:15-31 as smx_uniform_init), and fixed-pattern ADC maps in
batches (_epifm.py:926-963).  Statistical checks against the laws the
reference's draws follow."""
import math

import numpy
import pytest
import torch
from scipy import stats

from conftest import product_config

pytestmark = pytest.mark.gpu


def _binom_ok(frac, p, n, z=5.0):
    return abs(frac - p) <= z * math.sqrt(p * (1 - p) / n) + 1e-4


def test_philox_state_transitions_and_per_state_D(gpu):
    """Two states with rates k01, k10: one-step switch fractions equal
    1 - exp(-k dt) (sampling2.py:56-60), the long-run fraction of state 1 is
    that chain's stationary p01 / (p01 + p10), and each state's increments
    have variance 2 D_s dt."""
    import scopyon_amd
    k01, k10, dt, nsteps = 3.0, 1.0, 0.05, 60
    D = [0.02, 0.2]
    trans = numpy.array([[0.0, k01], [k10, 0.0]])
    t = numpy.arange(nsteps + 1) * dt
    n0, n1 = 30000, 10000
    r = scopyon_amd.sample(t, N=[n0, n1], lower=-100.0, upper=100.0, D=D, ndim=2,
                           transmat=trans, rng=scopyon_amd.PhiloxRNG(5))
    a = numpy.stack(r)                                  # (T+1, n, ndim + 2)
    s = a[:, :, 2].astype(int)
    assert (s[0, :n0] == 0).all() and (s[0, n0:] == 1).all()
    assert numpy.array_equal(a[:, :, 3], numpy.tile(numpy.arange(n0 + n1), (nsteps + 1, 1)))
    p01, p10 = 1 - math.exp(-k01 * dt), 1 - math.exp(-k10 * dt)
    from0 = s[0] == 0
    assert _binom_ok((s[1][from0] == 1).mean(), p01, from0.sum())
    assert _binom_ok((s[1][~from0] == 0).mean(), p10, (~from0).sum())
    # stationary occupation of the discrete chain the reference steps
    # (p01 / (p01 + p10), not the continuous-time k01 / (k01 + k10)); it
    # mixes in ~1 / (p01 + p10) = 5 steps
    late = (s[-20:] == 1).mean()
    assert abs(late - p01 / (p01 + p10)) < 0.006
    # a step moves with the state before the transition (move, then switch)
    inc = a[1:, :, :2] - a[:-1, :, :2]
    for state, Ds in enumerate(D):
        sel = s[:-1] == state
        z = inc[sel].ravel() / math.sqrt(2 * Ds * dt)
        assert stats.kstest(z, 'norm').pvalue > 1e-4, state


def test_philox_per_axis_D(gpu):
    """sample_inputs with a per-axis D (sampling.py:115-118): MSD per axis
    grows as 2 D_d t."""
    import scopyon_amd
    D = [0.1e-12, 0.5e-12, 2.0e-12]
    dt, nsteps = 0.01, 30
    t = numpy.arange(nsteps + 1) * dt
    r = scopyon_amd.sample_inputs(t, N=20000, lower=-1e-3, upper=1e-3, D=D, ndim=3,
                                  rng=scopyon_amd.PhiloxRNG(11))
    x = numpy.stack([p[:, :3] for _, p in r])
    msd = ((x - x[0]) ** 2).mean(axis=1)                 # (T+1, 3)
    for d in range(3):
        slope = numpy.polyfit(t, msd[:, d], 1)[0]
        assert abs(slope / (2 * D[d]) - 1) < 0.05, (d, slope)


def test_uniform_init(gpu):
    """smx_uniform_init: every axis uniform on [lower, upper) (KS), a
    degenerate axis constant, reproducible for a seed, different per seq."""
    from scopyon_amd import _device, _lib
    n = 200000
    xs = [torch.empty(n, dtype=torch.float64, device='cuda') for _ in range(3)]
    lo, hi = [-2.0, 5.0, 1.5], [3.0, 5.0, 4.0]

    def draw(seq):
        _lib.call('smx_uniform_init', n, 3, xs[0].data_ptr(), xs[1].data_ptr(), xs[2].data_ptr(),
                  _lib.host_doubles(lo), _lib.host_doubles(hi), 77, seq, 0, _device.stream())
        return [x.cpu().numpy() for x in xs]
    a = draw(0)
    for d in (0, 2):
        assert (a[d] >= lo[d]).all() and (a[d] < hi[d]).all()
        assert stats.kstest((a[d] - lo[d]) / (hi[d] - lo[d]), 'uniform').pvalue > 1e-4
    assert (a[1] == 5.0).all()
    b = draw(0)
    c = draw(1)
    assert numpy.array_equal(a[0], b[0]) and not numpy.array_equal(a[0], c[0])
    # x and z of one particle come from different words: uncorrelated
    assert abs(numpy.corrcoef(a[0], a[2])[0, 1]) < 0.01


def test_batched_fpn_maps(gpu):
    """Fixed-pattern ADC offsets per FOV (fpn 'pixel' and 'column'): the
    maps follow the reference's recipe at each FOV's size from one
    RandomState(seed), a frame without light and readout noise reads its
    offset map exactly, and the fused detector equals the separate one."""
    from scopyon_amd.batched import BatchedSimulator
    shapes = [(64, 64), (32, 48)]
    for kind in ('pixel', 'column'):
        print('batched fpn', kind, flush=True)
        ov = {'detector': {'image_size': [64, 64], 'exposure_time': 0.033,
                           'readout_noise': 0.0},
              'magnification': 100.0,
              'analog_to_digital_converter': {'type': kind, 'count': 30.0},
              'effects': {'photo_bleaching': {'switch': False}, 'background': {'switch': False}}}
        dark = BatchedSimulator(product_config(ov), shapes, [0, 0], dt=0.033, seed=3,
                                adc_dtype=torch.float64)
        rng = numpy.random.RandomState(3)
        want = []
        for nw, nh in shapes:
            if kind == 'pixel':
                off = numpy.rint(rng.normal(2000.0, 30.0, nw * nh))
            else:
                col = rng.normal(2000.0, 30.0, nh)
                off = numpy.rint(numpy.tile(col, (1, nw)).reshape(nh * nw))
            want.append(off)
        want = numpy.concatenate(want)
        assert numpy.array_equal(dark.offset_map.cpu().numpy(), want)
        dark.step()
        torch.cuda.synchronize()
        adc = dark.adc.cpu().numpy()
        assert numpy.array_equal(adc, numpy.float32(want).astype(float))
        kw = dict(dt=0.033, seed=9, tile_side=32)
        lit = {k: BatchedSimulator(product_config(ov), shapes, [80, 40], fuse_detector=f, **kw)
               for k, f in (('sep', False), ('fus', True))}
        for sim in lit.values():
            sim.step()
        torch.cuda.synchronize()
        assert torch.equal(lit['sep'].adc, lit['fus'].adc)
