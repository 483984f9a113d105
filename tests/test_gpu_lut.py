"""LUT mode (radial x depth lookup table, 256 depth slices x 4096 radial
bins, pixel integrals at 33 x 33 sub-pixel phases) on the 3D scale-up
geometry (66.4 nm pixels): device tables and frames against the float64
restatement (oracle/lut.py), its error against the continuous model and its
deviation from the 1 nm parity path, the fused detector, tile sides."""
import numpy
import pytest
import torch

from oracle import epifm as O
from oracle import lut as OL
from conftest import oracle_params, product_config

pytestmark = pytest.mark.gpu

LAM = 5.780000000000062e-07
PL = 16e-6 / 241.0
OV = {'detector': {'image_size': [48, 40], 'exposure_time': 0.01}, 'magnification': 241.0,
      'light_source': {'angle': 0.0}, 'effects': {'photo_bleaching': {'switch': False}}}


@pytest.fixture(scope='module')
def oracle():
    return OL.LutOracle(LAM, PL)


def _sim(shapes, counts, **kw):
    from scopyon_amd.batched import BatchedSimulator
    return BatchedSimulator(product_config(OV), shapes, counts, dt=0.01, psf='lut',
                            depth_range=(-1e-6, 1e-6), **kw)


def test_lut_tables_match_oracle(gpu, oracle):
    print('lut: tables', flush=True)
    sim = _sim([(32, 32)], [1], seed=1)
    t = sim.tables
    assert (t.nz, t.nb, t.lut.q, t.m) == (256, 4096, 32, oracle.m)
    rad = t.radial.cpu().numpy()
    for s in (0, 77, 255):
        ref = oracle.profile(s)
        assert numpy.abs(rad[s] - ref).max() <= 1e-12 * ref.max(), s
    K = t.K.view(t.nz, t.lut.q + 1, t.lut.q + 1, t.m, t.m)
    sums = []
    for s, qu, qv in ((0, 0, 0), (0, 17, 5), (77, 32, 32), (255, 3, 31), (128, 16, 0)):
        ref = oracle.block(s, qu, qv)
        got = K[s, qu, qv].cpu().numpy().astype(float)
        numpy.testing.assert_allclose(got, ref, rtol=2e-6, atol=1e-6 * ref.max(),
                                      err_msg=str((s, qu, qv)))
        if s == 0:
            sums.append(got.sum())
    # every phase's footprint holds the whole kernel: same photon total
    assert abs(sums[0] - sums[1]) <= 1e-6 * sums[0]
    # the splat's cell form: per pixel the four phases around cell (cu, cv)
    # in tap order (cu, cv), (cu, cv + 1), (cu + 1, cv), (cu + 1, cv + 1)
    q = t.lut.q
    C = t.cells.view(t.nz, q, q, t.m, t.m, 4)
    for s, cu, cv in ((0, 0, 0), (77, 31, 31), (255, 12, 30)):
        got = C[s, cu, cv].cpu().numpy()
        for k, (a, b) in enumerate(((0, 0), (0, 1), (1, 0), (1, 1))):
            assert numpy.array_equal(got[..., k], K[s, cu + a, cv + b].cpu().numpy()), (s, cu, cv, k)


def _emitters(sim, rng, depths):
    u, v = sim.u.cpu().numpy(), sim.v.cpu().numpy()
    d = rng.choice(depths, sim.n)
    sim.set_positions(d, u, v)
    sim.weights()
    torch.cuda.synchronize()
    return d, u, v, sim.w.cpu().numpy()


@pytest.mark.parametrize('ts', [16, 32])
def test_lut_frames_match_oracle(gpu, oracle, ts):
    print('lut: frames vs oracle, tile side {}'.format(ts), flush=True)
    shapes, counts = [(48, 40), (32, 32)], [25, 15]
    sim = _sim(shapes, counts, seed=4, tile_side=ts)
    rng = numpy.random.RandomState(3)
    d, u, v, w = _emitters(sim, rng, [-3.3e-9, 120e-9, -457e-9, 999e-9, 2e-6])
    sim.splat()
    torch.cuda.synchronize()
    assert sim.status() & ~1 == 0
    p0 = sim.configs.detector_focal_point
    for f, shape in enumerate(shapes):
        a, b = sim.emitter_offsets[f], sim.emitter_offsets[f + 1]
        ref = OL.splat_lut(oracle, shape, d[a:b], u[a:b], v[a:b], w[a:b], p0[0], p0[1], p0[2])
        got = sim.frame(f, 'photons').cpu().numpy()
        numpy.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-6 * ref.max(), err_msg=str(f))


def test_lut_error_budget(gpu, oracle):
    """LUT frames against the continuous model (the kernel integrated over
    the exact pixel at the exact position): at the depth of the emitter's
    slice, within 5e-4 of the frame maximum (bilinear phase interpolation,
    float32 tables); at the emitter's exact depth within 5e-3 (slices are
    3.9 nm apart, so up to ~2.5 nm of defocus with the 1 nm key).  Against
    the 1 nm parity path (float64 oracle), which rounds each pixel edge to
    its 1 nm cell grid per emitter (up to ~1.5% of a pixel's light, not
    reproduced here): within 3% of the frame maximum."""
    print('lut: error budget', flush=True)
    shape = (40, 40)
    sim = _sim([shape], [6], seed=8, tile_side=32)
    rng = numpy.random.RandomState(12)
    u = rng.uniform(-12 * PL, 12 * PL, 6)
    v = rng.uniform(-12 * PL, 12 * PL, 6)
    d = numpy.array([0.0, 37.3e-9, -210.6e-9, 480.2e-9, -733.9e-9, 998.1e-9])
    sim.set_positions(d, u, v)
    sim.weights()
    sim.splat()
    torch.cuda.synchronize()
    got = sim.frame(0, 'photons').cpu().numpy()
    w = sim.w.cpu().numpy()
    half = shape[0] * PL * 0.5
    conts = []
    for depth_of in (lambda i: oracle.depths[OL.slice_of_depth(d[i], 0.0, oracle.depth_cutoff,
                                                                oracle.nz)],
                     lambda i: d[i]):
        cont = numpy.zeros(shape)
        for i in range(6):
            i0, j0, blk = OL.pixel_integral_exact(oracle, depth_of(i), (half + u[i]) / PL,
                                                  (half + v[i]) / PL)
            a0, b0 = max(i0, 0), max(j0, 0)
            a1, b1 = min(i0 + oracle.m, shape[0]), min(j0 + oracle.m, shape[1])
            cont[a0:a1, b0:b1] += w[i] * blk[a0 - i0:a1 - i0, b0 - j0:b1 - j0]
        conts.append(cont)
    err_s = numpy.abs(got - conts[0]).max() / conts[0].max()
    err_c = numpy.abs(got - conts[1]).max() / conts[1].max()
    p = oracle_params(OV)
    parts = numpy.column_stack([d, u, v, numpy.arange(6), numpy.ones(6)])
    par, _, _ = O.molecule_plane(p, O.PSFCache(p), parts, shape, numpy.zeros(3), 0.01, {}, None,
                                 None)
    err_p = numpy.abs(got - par).max() / par.max()
    rms_p = numpy.sqrt(((got - par) ** 2).mean()) / par.max()
    print('lut vs continuous at the slice depth: max {:.2e} of peak; at the exact depth: '
          'max {:.2e}; vs 1 nm parity: max {:.2e}, rms {:.2e}'.format(err_s, err_c, err_p, rms_p),
          flush=True)
    assert err_s < 5e-4
    assert err_c < 5e-3
    assert err_p < 3e-2


@pytest.mark.parametrize('ts,q', [(16, 16), (32, 16), (64, 16), (32, 32), (64, 32)])
def test_lut_fused_detector_and_tile_sides(gpu, ts, q):
    """Fused LUT splat + EMCCD equals the separate passes bit for bit; frames
    do not depend on the tile shape."""
    print('lut: fused detector, tile side {} quadrant {}'.format(ts, q), flush=True)
    shapes, counts = [(64, 64), (40, 56)], [60, 30]
    kw = dict(seed=6, tile_side=ts, quadrant=q, keep_expectation=True)
    sep = _sim(shapes, counts, fuse_detector=False, **kw)
    fus = _sim(shapes, counts, fuse_detector=True, keep_photons=True, **kw)
    assert fus.fused and not sep.fused
    for s in (sep, fus):
        s.weights()
    sep.splat()
    sep.emccd()
    fus.splat_detect()
    torch.cuda.synchronize()
    assert numpy.array_equal(sep.adc.cpu().numpy(), fus.adc.cpu().numpy())
    assert numpy.array_equal(sep.photons.cpu().numpy(), fus.photons.cpu().numpy())
    ref = _sim(shapes, counts, seed=6, tile_side=32, fuse_detector=False)
    ref.weights()
    ref.splat()
    assert numpy.array_equal(ref.photons.cpu().numpy().view(numpy.uint64),
                             sep.photons.cpu().numpy().view(numpy.uint64))
