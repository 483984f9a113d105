"""Splat + EMCCD + ADC in one tile pass (smx_splat_detect_ts) against the two
separate passes (smx_splat_expected_ts, then smx_emccd_philox): the fused
epilogue draws every pixel's numbers from the same Philox stream with the
same code, so ADC counts, expectations and photons are bit-identical, for
tiles of 16, 32 and 64 pixels, uneven frames and several steps."""
import ctypes

import numpy
import pytest
import torch

from conftest import product_config

pytestmark = pytest.mark.gpu

SHAPES = [(64, 64), (48, 80), (128, 128), (16, 16), (200, 72)]
COUNTS = [50, 10, 400, 3, 120]


def _cfg():
    return product_config({'detector': {'image_size': [64, 64], 'exposure_time': 0.033},
                           'magnification': 100.0,
                           'effects': {'photo_bleaching': {'switch': False}}})


def _pair(ts, precision='folded', adc_dtype=torch.float32, keep_photons=True, seed=3, q=16):
    from scopyon_amd.batched import BatchedSimulator
    kw = dict(dt=0.033, seed=seed, table_precision=precision, tile_side=ts, quadrant=q,
              adc_dtype=adc_dtype, keep_expectation=True)
    sep = BatchedSimulator(_cfg(), SHAPES, COUNTS, fuse_detector=False, **kw)
    fus = BatchedSimulator(_cfg(), SHAPES, COUNTS, fuse_detector=True, keep_photons=keep_photons,
                           **kw)
    assert not sep.fused and fus.fused
    return sep, fus


def _bits(t):
    a = t.cpu().numpy()
    return a.view(numpy.uint64 if a.dtype == numpy.float64 else numpy.uint32)


@pytest.mark.parametrize('ts,q', [(16, 16), (32, 16), (64, 16), (32, 32), (64, 32)])
@pytest.mark.parametrize('precision', ['folded', 'float64'])
def test_fused_equals_separate_passes(gpu, ts, q, precision):
    print('fused detector: tile side {} quadrant {} table {}'.format(ts, q, precision), flush=True)
    sep, fus = _pair(ts, precision, q=q)
    for sim in (sep, fus):
        sim.photons.fill_(-1.0)
        sim.adc.fill_(-1.0)
        sim.weights()
    sep.splat()
    sep.emccd()
    fus.splat_detect()
    torch.cuda.synchronize()
    assert sep.status() & ~1 == 0 and fus.status() & ~1 == 0
    assert (sep.adc >= 0).all() and sep.photons.max() > 0
    assert numpy.array_equal(_bits(fus.adc), _bits(sep.adc))
    assert numpy.array_equal(_bits(fus.E), _bits(sep.E))
    assert numpy.array_equal(_bits(fus.photons), _bits(sep.photons))


@pytest.mark.parametrize('ts', [16, 32])
def test_fused_steps_and_graph_replay(gpu, ts):
    """Several steps (Brownian move, weights, fused pass, frame counter), the
    fused simulator replayed from a HIP graph, against eager separate steps;
    no photon image is written in the fused steps."""
    print('fused detector: steps, tile side {}'.format(ts), flush=True)
    sep, fus = _pair(ts, keep_photons=False, adc_dtype=torch.float64, seed=7)
    fus.photons.fill_(-2.0)
    fus.capture()                                # one eager step, then the graph
    outs = []
    sep.step()
    outs.append((sep.adc.clone(), fus.adc.clone()))
    for _ in range(3):
        sep.step()
        fus.replay()
        outs.append((sep.adc.clone(), fus.adc.clone()))
    torch.cuda.synchronize()
    for a, b in outs:
        assert numpy.array_equal(_bits(a), _bits(b))
    assert not torch.equal(outs[0][0], outs[1][0])
    assert (fus.photons == -2.0).all()
    assert torch.equal(sep.u, fus.u) and torch.equal(sep.v, fus.v)


def test_fused_with_adc_maps(gpu):
    """Per-pixel gain and offset maps (fixed-pattern noise) through the C-ABI:
    the fused pass reads them like smx_emccd_philox."""
    from scopyon_amd import _device, _lib
    sep, fus = _pair(32, keep_photons=True, seed=13)
    n = sep.n_pix
    rng = numpy.random.RandomState(5)
    gain = torch.as_tensor(rng.uniform(10.0, 14.0, n), device='cuda')
    offset = torch.as_tensor(numpy.rint(rng.normal(2000.0, 30.0, n)), device='cuda')
    sep.weights()
    sep.splat()
    adc_a = torch.zeros(n, dtype=torch.float32, device='cuda')
    _lib.call('smx_emccd_philox', n, sep.photons.data_ptr(), gain.data_ptr(), offset.data_ptr(),
              sep.det, 99, 4, 0, None, None, adc_a.data_ptr(), 1, _device.stream())
    fus.weights()
    t = fus.tables
    adc_b = torch.zeros(n, dtype=torch.float32, device='cuda')
    _lib.call('smx_splat_detect_ts', fus.n, fus.u.data_ptr(), fus.v.data_ptr(), fus.w.data_ptr(),
              fus.key.data_ptr(), fus.fov_id.data_ptr(), fus.fovs_dev.data_ptr(), fus.nfov,
              fus.n_tiles, t.sat.data_ptr(), t.sat_elems, len(t.slot_of), t.ntab, t.period, 1e-9,
              1e-18, fus.capacity, 0, _device.ptr(fus.Ktab), fus.k_format, t.m, None,
              fus.ws.data_ptr(), fus.ws.numel(), fus.splatter.status.data_ptr(), _device.stream(),
              gain.data_ptr(), offset.data_ptr(), fus.det, 99, 4, 0, None, None, adc_b.data_ptr(),
              1, 32)
    torch.cuda.synchronize()
    assert numpy.array_equal(_bits(adc_a), _bits(adc_b))


def test_fused_refuses_what_it_cannot_do(gpu):
    from scopyon_amd import _device, _lib
    sep, fus = _pair(32)
    t = fus.tables
    args = [fus.n, fus.u.data_ptr(), fus.v.data_ptr(), fus.w.data_ptr(), fus.key.data_ptr(),
            fus.fov_id.data_ptr(), fus.fovs_dev.data_ptr(), fus.nfov, fus.n_tiles,
            t.sat.data_ptr(), t.sat_elems, len(t.slot_of), t.ntab, t.period, 1e-9, 1e-18,
            fus.capacity]
    tail = [fus.ws.data_ptr(), fus.ws.numel(), fus.splatter.status.data_ptr(), _device.stream(),
            None, None, fus.det, 1, 0, 0, None, None, fus.adc.data_ptr(), 1, 32]
    with pytest.raises(ValueError):       # accumulate = 1
        _lib.call('smx_splat_detect_ts', *args, 1, _device.ptr(fus.Ktab), fus.k_format, t.m,
                  None, *tail)
    with pytest.raises(ValueError):       # no phase table
        _lib.call('smx_splat_detect_ts', *args, 0, None, 0, t.m, None, *tail)
