"""Batched multi-FOV engine: per-FOV frames equal the oracle's single-frame
result, and HIP-graph replays equal eager steps."""
import numpy
import pytest
import torch

from oracle import epifm as O
from conftest import oracle_params, product_config

pytestmark = pytest.mark.gpu


def _cfg(shape=(64, 64)):
    return product_config({'detector': {'image_size': list(shape), 'exposure_time': 0.033},
                           'magnification': 100.0,
                           'effects': {'photo_bleaching': {'switch': False}}})


@pytest.mark.parametrize('precision,rtol', [('float64', 1e-11), ('float32', 1e-6),
                                            ('folded', 1e-6)])
def test_uneven_fovs_match_oracle(gpu, precision, rtol):
    from scopyon_amd.batched import BatchedSimulator
    shapes = [(64, 64), (48, 80), (128, 128), (16, 16)]
    counts = [50, 10, 300, 0]
    sim = BatchedSimulator(_cfg(), shapes, counts, dt=0.033, seed=3, table_precision=precision)
    sim.photons.fill_(7.0)      # form() overwrites every pixel
    sim.form()
    assert sim.status() & ~1 == 0
    u, v = sim.u.cpu().numpy(), sim.v.cpu().numpy()
    for f, (shape, k) in enumerate(zip(shapes, counts)):
        a, b = sim.emitter_offsets[f], sim.emitter_offsets[f + 1]
        p = oracle_params({'detector': {'image_size': list(shape)}, 'magnification': 100.0})
        parts = numpy.column_stack([numpy.zeros(k), u[a:b], v[a:b], numpy.arange(k), numpy.ones(k)])
        ref, _, _ = O.molecule_plane(p, O.PSFCache(p), parts, shape, numpy.zeros(3), 0.033, {},
                                     None, None)
        got = sim.frame(f, 'photons').cpu().numpy()
        if k == 0:
            assert not got.any()
        else:
            numpy.testing.assert_allclose(got, ref, rtol=rtol, atol=rtol * 1e-2 * ref.max())


def test_many_fovs_span_the_binning_window(gpu):
    """70 FOVs of 256^2 (64 tiles each) with 15 emitters each: the first block
    of 1024 emitters spans more tiles than its LDS counting window (4096), so
    tile counts and slot claims mix the LDS-aggregated and the direct global
    path; every FOV still equals the oracle."""
    from scopyon_amd.batched import BatchedSimulator
    shape, nf, k = (256, 256), 70, 15
    sim = BatchedSimulator(_cfg(shape), [shape] * nf, [k] * nf, dt=0.033, seed=5, tile_side=32)
    assert sim.n_tiles > 4096 and nf * k > 1024
    sim.form()
    assert sim.status() & ~1 == 0
    u, v = sim.u.cpu().numpy(), sim.v.cpu().numpy()
    p = oracle_params({'detector': {'image_size': list(shape)}, 'magnification': 100.0})
    cache = O.PSFCache(p)
    for f in range(nf):
        a, b = sim.emitter_offsets[f], sim.emitter_offsets[f + 1]
        parts = numpy.column_stack([numpy.zeros(k), u[a:b], v[a:b], numpy.arange(k), numpy.ones(k)])
        ref, _, _ = O.molecule_plane(p, cache, parts, shape, numpy.zeros(3), 0.033, {}, None, None)
        got = sim.frame(f, 'photons').cpu().numpy()
        numpy.testing.assert_allclose(got, ref, rtol=1e-6, atol=1e-8 * ref.max(), err_msg=str(f))


@pytest.mark.parametrize('precision', ['float64', 'folded'])
def test_tile_sides_give_identical_frames(gpu, precision):
    """Tiles of 16, 32 and 64 pixels with 16 x 16 quadrants (1, 4 and 16
    waves per tile) and tiles of 32 and 64 pixels with 32 x 32 quadrants (1
    and 4 waves) sum every pixel's emitters in emitter order, so the frames
    are bit-identical; the autotuner times all five and keeps one."""
    from scopyon_amd.batched import BatchedSimulator
    shapes = [(64, 64), (48, 80), (128, 128), (16, 16), (200, 72)]
    counts = [50, 10, 400, 3, 120]
    frames = {}
    for ts, q in BatchedSimulator.TILE_SHAPES:
        print('tile shape {} / {}'.format(ts, q), flush=True)
        sim = BatchedSimulator(_cfg(), shapes, counts, dt=0.033, seed=3, table_precision=precision,
                               tile_side=ts, quadrant=q)
        assert (sim.tile_side, sim.quadrant) == (ts, q)
        sim.form()
        assert sim.status() & ~1 == 0
        frames[ts, q] = sim.photons.cpu().numpy()
    assert frames[32, 16].any()
    for k in frames:
        assert numpy.array_equal(frames[k].view(numpy.uint64),
                                 frames[32, 16].view(numpy.uint64)), k
    auto = BatchedSimulator(_cfg(), shapes, counts, dt=0.033, seed=3, table_precision=precision)
    assert (auto.tile_side, auto.quadrant) in BatchedSimulator.TILE_SHAPES
    assert sorted(auto.tile_timings) == sorted(['16', '32', '64', '32/q32', '64/q32'])
    auto.form()
    assert numpy.array_equal(auto.photons.cpu().numpy().view(numpy.uint64),
                             frames[32, 16].view(numpy.uint64))
    with pytest.raises(ValueError):
        BatchedSimulator(_cfg(), shapes, counts, dt=0.033, seed=3, tile_side=24)
    with pytest.raises(ValueError):
        BatchedSimulator(_cfg(), shapes, counts, dt=0.033, seed=3, tile_side=16, quadrant=32)


# EmitterRec as the phase-table path stores it at the start of the workspace
_REC = numpy.dtype({'names': ['lo_u', 'lo_v', 'n_u', 'n_v', 'su', 'sv', 'flags', 'w', 'table_off'],
                    'formats': ['<i4', '<i4', '<i2', '<i2', '<i2', '<i2', '<i4', '<f8', '<i8'],
                    'offsets': [0, 4, 8, 10, 12, 14, 16, 24, 32], 'itemsize': 40})


@pytest.mark.parametrize('precision,rtol', [('float64', 1e-11), ('folded', 1e-6)])
@pytest.mark.parametrize('tile_side', [16, 32])
def test_irregular_emitters_match_oracle(gpu, monkeypatch, precision, rtol, tile_side):
    """Emitters on the 1 nm grid of the PSF table put pixel edges exactly on
    table cells, where float64 rounding of a boundary quotient can leave the
    regular phase grid: such emitters are flagged and integrated through the
    SAT, one part at a time, in list order with the regular ones.  Half the
    emitters here sit on the grid (nearly all flagged), half are random; the
    frame equals the oracle, and the 64-bit table-address form of the tile
    pass gives the same bits as the 32-bit one."""
    from scopyon_amd.batched import BatchedSimulator
    shape, k = (64, 64), 300
    sim = BatchedSimulator(_cfg(shape), [shape], [k], dt=0.033, seed=9, tile_side=tile_side,
                           table_precision=precision)
    rng = numpy.random.RandomState(4)
    pos = rng.uniform(-5e-6, 5e-6, size=(k, 2))
    pos[::2] = rng.randint(-5000, 5000, size=(k // 2, 2)) * 1e-9
    sim.u.copy_(torch.from_numpy(pos[:, 0].copy()))
    sim.v.copy_(torch.from_numpy(pos[:, 1].copy()))
    sim.form()
    assert sim.status() & ~1 == 0
    rec = numpy.frombuffer(sim.ws[:k * _REC.itemsize].cpu().numpy().tobytes(), dtype=_REC)
    flagged = rec['flags'] == 1
    assert flagged[::2].sum() > k // 4 and not flagged[1::2].any()
    p = oracle_params({'detector': {'image_size': list(shape)}, 'magnification': 100.0})
    parts = numpy.column_stack([numpy.zeros(k), pos[:, 0], pos[:, 1], numpy.arange(k), numpy.ones(k)])
    ref, _, _ = O.molecule_plane(p, O.PSFCache(p), parts, shape, numpy.zeros(3), 0.033, {}, None, None)
    got = sim.frame(0, 'photons').cpu().numpy()
    numpy.testing.assert_allclose(got, ref, rtol=rtol, atol=rtol * 1e-2 * ref.max())
    first = sim.photons.cpu().numpy()
    monkeypatch.setenv('SMX_SPLAT_ADDR64', '1')
    sim.form()
    assert numpy.array_equal(sim.photons.cpu().numpy().view(numpy.uint64), first.view(numpy.uint64))


def test_long_tile_lists_match_oracle(gpu):
    """6,000 emitters on one 64 x 64 frame give tile lists longer than the
    tile pass sorts in LDS (1,024 entries for 16 px tiles, 2,048 for 32 and
    64 px): they are sorted after binning, so every pixel still sums its
    emitters in index order -- frames equal the oracle to rounding, are
    bit-identical across tile sides and from run to run, and no status bit
    is set.  With 10,000 emitters a 64 px tile's list exceeds the 8,192
    entries that sort holds: status bit 1, values still equal the oracle."""
    from scopyon_amd.batched import BatchedSimulator
    shape = (64, 64)
    p = oracle_params({'detector': {'image_size': list(shape)}, 'magnification': 100.0})
    for k, sides in ((6000, (16, 32, 64)), (10000, (64,))):
        print('long lists: {} emitters'.format(k), flush=True)
        ref = first = None
        for ts in sides:
            sim = BatchedSimulator(_cfg(shape), [shape], [k], dt=0.033, seed=17, tile_side=ts,
                                   table_precision='float64')
            sim.form()
            if k <= 8192:
                assert sim.status() == 0
            else:
                assert sim.status() & 1 and sim.status() & ~1 == 0
            if ref is None:
                u, v = sim.u.cpu().numpy(), sim.v.cpu().numpy()
                parts = numpy.column_stack([numpy.zeros(k), u, v, numpy.arange(k), numpy.ones(k)])
                ref, _, _ = O.molecule_plane(p, O.PSFCache(p), parts, shape, numpy.zeros(3), 0.033,
                                             {}, None, None)
            got = sim.frame(0, 'photons').cpu().numpy()
            numpy.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-12 * ref.max(),
                                          err_msg=str(ts))
            if k <= 8192:
                if first is None:
                    first = got
                assert numpy.array_equal(got.view(numpy.uint64), first.view(numpy.uint64)), ts
                sim.form()
                again = sim.frame(0, 'photons').cpu().numpy()
                assert numpy.array_equal(again.view(numpy.uint64), got.view(numpy.uint64)), ts


def test_mid_length_tile_lists_match_oracle(gpu):
    """700 and 1,500 emitters on one 64 x 64 frame: a 32 px tile's list then
    holds ~360 or ~780 entries and a 64 px tile's all of them, beyond one
    entry per thread of the block, so the tile pass sorts them with the
    merge-rank sort (wave-sorted runs ranked against each other) -- frames
    equal the oracle to rounding and are bit-identical across tile sides and
    from run to run."""
    from scopyon_amd.batched import BatchedSimulator
    shape = (64, 64)
    p = oracle_params({'detector': {'image_size': list(shape)}, 'magnification': 100.0})
    for k in (700, 1500):
        print('mid-length lists: {} emitters'.format(k), flush=True)
        ref = first = None
        for ts in (16, 32, 64):
            sim = BatchedSimulator(_cfg(shape), [shape], [k], dt=0.033, seed=23, tile_side=ts,
                                   table_precision='float64')
            sim.form()
            assert sim.status() == 0
            if ref is None:
                u, v = sim.u.cpu().numpy(), sim.v.cpu().numpy()
                parts = numpy.column_stack([numpy.zeros(k), u, v, numpy.arange(k), numpy.ones(k)])
                ref, _, _ = O.molecule_plane(p, O.PSFCache(p), parts, shape, numpy.zeros(3), 0.033,
                                             {}, None, None)
            got = sim.frame(0, 'photons').cpu().numpy()
            numpy.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-12 * ref.max(),
                                          err_msg=str(ts))
            if first is None:
                first = got
            assert numpy.array_equal(got.view(numpy.uint64), first.view(numpy.uint64)), ts
            sim.form()
            again = sim.frame(0, 'photons').cpu().numpy()
            assert numpy.array_equal(again.view(numpy.uint64), got.view(numpy.uint64)), ts


@pytest.mark.parametrize('precision', ['float64', 'float32', 'folded'])
def test_table_address_forms_agree(gpu, monkeypatch, precision):
    """The tile pass addresses the phase table with 32-bit byte offsets from
    its base when the table spans less than 4 GiB, else with 64-bit addresses
    per part; both forms give the same bits (tiles of 16, 32, 64)."""
    from scopyon_amd.batched import BatchedSimulator
    shapes = [(64, 64), (48, 80), (128, 128), (200, 72)]
    counts = [50, 10, 900, 120]
    for ts in (16, 32, 64):
        monkeypatch.delenv('SMX_SPLAT_ADDR64', raising=False)
        sim = BatchedSimulator(_cfg(), shapes, counts, dt=0.033, seed=21, table_precision=precision,
                               tile_side=ts)
        sim.form()
        assert sim.status() & ~1 == 0
        a = sim.photons.cpu().numpy()
        monkeypatch.setenv('SMX_SPLAT_ADDR64', '1')
        sim.form()
        assert numpy.array_equal(sim.photons.cpu().numpy().view(numpy.uint64), a.view(numpy.uint64)), ts


def test_graph_replay_equals_eager(gpu):
    from scopyon_amd.batched import BatchedSimulator
    shapes = [(64, 64), (32, 96)]
    a = BatchedSimulator(_cfg(), shapes, [200, 80], dt=0.033, seed=11)
    b = BatchedSimulator(_cfg(), shapes, [200, 80], dt=0.033, seed=11)
    outs = []
    for _ in range(4):
        a.step()
        outs.append(a.adc.clone())
    b.capture()            # capture() runs one eager warm-up step first
    got = [b.adc.clone()]
    for _ in range(3):
        b.replay()
        got.append(b.adc.clone())
    torch.cuda.synchronize()
    for x, y in zip(outs, got):
        assert torch.equal(x, y)
    assert not torch.equal(outs[0], outs[1])     # counters advance between steps
    assert torch.equal(a.u, b.u) and torch.equal(a.v, b.v)


def test_multi_step_graph_equals_eager(gpu):
    """Three steps captured in one HIP graph, replayed twice, equal six eager
    steps (ADC of the last step of each graph, positions)."""
    from scopyon_amd.batched import BatchedSimulator
    shapes = [(64, 64), (32, 96)]
    a = BatchedSimulator(_cfg(), shapes, [200, 80], dt=0.033, seed=19)
    b = BatchedSimulator(_cfg(), shapes, [200, 80], dt=0.033, seed=19)
    outs = []
    for k in range(7):
        a.step()
        if k in (0, 3, 6):
            outs.append(a.adc.clone())
    b.capture(steps=3)     # one eager warm-up step, then the 3-step graph
    got = [b.adc.clone()]
    for _ in range(2):
        b.replay(3)
        got.append(b.adc.clone())
    torch.cuda.synchronize()
    for x, y in zip(outs, got):
        assert torch.equal(x, y)
    assert torch.equal(a.u, b.u) and torch.equal(a.v, b.v)
    with pytest.raises(ValueError):
        b.replay(2)


def test_reflecting_walls_hold(gpu):
    from scopyon_amd.batched import BatchedSimulator
    sim = BatchedSimulator(_cfg(), [(32, 32)], 500, D=5e-12, dt=0.033, seed=2)
    for _ in range(20):
        sim.move()
        torch.cuda.synchronize()
        sim.counter += 1
    half = 0.5 * 32 * 1.6e-7
    u, v = sim.u.cpu().numpy(), sim.v.cpu().numpy()
    assert (numpy.abs(u) <= half).all() and (numpy.abs(v) <= half).all()


def test_3d_batch_epi_matches_oracle(gpu):
    """3D batch at magnification 241 (66 nm pixels: the general summed-area
    path), EPI illumination, depth-dependent Born-Wolf tables for every key the
    depth range reaches; noise-free frames vs the float64 oracle."""
    from scopyon_amd.batched import BatchedSimulator
    ov = {'detector': {'image_size': [48, 40], 'exposure_time': 0.01}, 'magnification': 241.0,
          'light_source': {'angle': 0.0},
          'effects': {'photo_bleaching': {'switch': False}}}
    shapes = [(48, 40), (32, 32)]
    counts = [25, 15]
    sim = BatchedSimulator(product_config(ov), shapes, counts, dt=0.01, seed=4,
                           depth_range=(-4e-9, 4e-9), D=numpy.full(40, 0.5e-12))
    rng = numpy.random.RandomState(3)
    depth = rng.choice([-3.3e-9, -1.4e-9, 0.2e-9, 2.6e-9], 40)
    u, v = sim.u.cpu().numpy(), sim.v.cpu().numpy()
    sim.set_positions(depth, u, v)
    sim.form()
    assert sim.status() & ~1 == 0
    for f, (shape, k) in enumerate(zip(shapes, counts)):
        a, b = sim.emitter_offsets[f], sim.emitter_offsets[f + 1]
        p = oracle_params({'detector': {'image_size': list(shape), 'exposure_time': 0.01},
                           'magnification': 241.0, 'light_source': {'angle': 0.0}})
        parts = numpy.column_stack([depth[a:b], u[a:b], v[a:b], numpy.arange(k), numpy.ones(k)])
        ref, _, _ = O.molecule_plane(p, O.PSFCache(p), parts, shape, numpy.zeros(3), 0.01, {},
                                     None, None)
        got = sim.frame(f, 'photons').cpu().numpy()
        numpy.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-13 * ref.max())


def test_3d_batch_full_depth_range_matches_oracle(gpu):
    """The 3D scale-up's parity form at its full depth range: +-1 um with 1 nm
    keys builds every key up to the depth cutoff (1002 float64 summed-area
    tables, 32 GB); emitters spread over the whole range, the cutoff key
    included, against the float64 oracle (general summed-area path, EPI)."""
    from scopyon_amd.batched import BatchedSimulator
    ov = {'detector': {'image_size': [40, 36], 'exposure_time': 0.01}, 'magnification': 241.0,
          'light_source': {'angle': 0.0},
          'effects': {'photo_bleaching': {'switch': False}}}
    depths = numpy.array([-999.6e-9, -987.4e-9, -512.3e-9, -100.2e-9, 0.7e-9, 333.3e-9,
                          777.7e-9, 999.9e-9, 1000.4e-9])
    k = len(depths)
    print('3D parity tables over +-1 um', flush=True)
    sim = BatchedSimulator(product_config(ov), [(40, 36)], [k], dt=0.01, seed=6,
                           depth_range=(-1e-6, 1e-6), D=numpy.full(k, 0.5e-12), psf='table')
    assert len(sim.tables.slot_of) >= 1002
    rng = numpy.random.RandomState(8)
    pl = 16e-6 / 241.0
    u, v = rng.uniform(-14 * pl, 14 * pl, k), rng.uniform(-12 * pl, 12 * pl, k)
    sim.set_positions(depths, u, v)
    sim.form()
    assert sim.status() & ~1 == 0
    p = oracle_params({'detector': {'image_size': [40, 36], 'exposure_time': 0.01},
                       'magnification': 241.0, 'light_source': {'angle': 0.0}})
    parts = numpy.column_stack([depths, u, v, numpy.arange(k), numpy.ones(k)])
    ref, _, _ = O.molecule_plane(p, O.PSFCache(p), parts, (40, 36), numpy.zeros(3), 0.01, {},
                                 None, None)
    got = sim.frame(0, 'photons').cpu().numpy()
    numpy.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-13 * ref.max())
    del sim
    torch.cuda.empty_cache()


@pytest.mark.parametrize('kind', ['CCD', 'CMOS'])
def test_batched_ccd_cmos_steps(gpu, kind):
    from scopyon_amd.batched import BatchedSimulator
    ov = {'detector': {'type': kind, 'image_size': [64, 64], 'exposure_time': 0.033},
          'magnification': 100.0, 'effects': {'photo_bleaching': {'switch': False}}}
    sim = BatchedSimulator(product_config(ov), [(64, 64), (32, 48)], [100, 40], dt=0.033,
                           seed=5)
    sim.capture()
    sim.replay(3)
    adc = sim.adc.cpu().numpy()
    assert numpy.isfinite(adc).all() and (adc >= 0).all()
    assert adc.std() > 0


def test_empty_fovs(gpu):
    """FOVs without emitters (and a batch without any) form frames of zeros,
    detect them (readout noise, ADC offset) and replay from a HIP graph; the
    FOV between them equals the oracle."""
    from scopyon_amd.batched import BatchedSimulator
    shapes, counts = [(64, 64), (48, 80), (32, 32)], [0, 25, 0]
    sim = BatchedSimulator(_cfg(), shapes, counts, dt=0.033, seed=41, tile_side=32,
                           table_precision='float64')
    sim.form()
    assert sim.status() & ~1 == 0
    assert not sim.frame(0, 'photons').cpu().numpy().any()
    assert not sim.frame(2, 'photons').cpu().numpy().any()
    u, v = sim.u.cpu().numpy(), sim.v.cpu().numpy()
    p = oracle_params({'detector': {'image_size': [48, 80]}, 'magnification': 100.0})
    parts = numpy.column_stack([numpy.zeros(25), u, v, numpy.arange(25), numpy.ones(25)])
    ref, _, _ = O.molecule_plane(p, O.PSFCache(p), parts, (48, 80), numpy.zeros(3), 0.033, {},
                                 None, None)
    numpy.testing.assert_allclose(sim.frame(1, 'photons').cpu().numpy(), ref, rtol=1e-11,
                                  atol=1e-13 * ref.max())
    sim.capture()
    sim.replay(2)
    adc = sim.adc.cpu().numpy()
    assert numpy.isfinite(adc).all() and (adc >= 0).all()
    none = BatchedSimulator(_cfg(), [(64, 64), (32, 48)], [0, 0], dt=0.033, seed=42)
    assert none.n == 0
    none.capture()
    none.replay(2)
    assert not none.photons.cpu().numpy().any() and none.status() & ~1 == 0
