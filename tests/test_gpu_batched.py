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
