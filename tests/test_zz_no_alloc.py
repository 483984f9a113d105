"""The allocation-free splat (smx_splat_expected_static: binning workspace in
the library's fixed device array) against the workspace form
(smx_splat_expected_ts) through the C-ABI.  Both run the same kernels on the
same sorted tile lists, so frames must be bit-identical; two calls in a row,
the second with fewer emitters and frames, show that nothing left in the
shared array from the first call leaks into the second."""
import numpy
import pytest
import torch

from conftest import product_config

pytestmark = pytest.mark.gpu


def _cfg(shape):
    return product_config({'detector': {'image_size': list(shape), 'exposure_time': 0.033},
                           'magnification': 100.0,
                           'effects': {'photo_bleaching': {'switch': False}}})


def _splat(sim, static, with_table=True, capacity=None):
    """One splat of `sim`'s emitters into a fresh frame buffer (filled with -1
    first, so an unwritten pixel shows); returns (frames, status word)."""
    from scopyon_amd import _device, _lib
    t = sim.tables
    out = torch.full((sim.n_pix,), -1.0, dtype=torch.float64, device=sim.device)
    status = torch.zeros(1, dtype=torch.int32, device=sim.device)
    K = _device.ptr(sim.Ktab) if with_table else None
    ts = sim.tile_side if with_table else 32
    cap = sim.capacity if capacity is None else capacity
    args = [sim.n, sim.u.data_ptr(), sim.v.data_ptr(), sim.w.data_ptr(), sim.key.data_ptr(),
            sim.fov_id.data_ptr(), sim.fovs_dev.data_ptr(), sim.nfov, sim.n_tiles,
            t.sat.data_ptr(), t.sat_elems, len(t.slot_of), t.ntab, t.period, 1e-9, 1e-18, cap, 0,
            K, sim.k_format if with_table else 0, t.m, out.data_ptr()]
    if static:
        _lib.call('smx_splat_expected_static', *args, status.data_ptr(), _device.stream(), ts)
    else:
        _lib.call('smx_splat_expected_ts', *args, sim.ws.data_ptr(), sim.ws.numel(),
                  status.data_ptr(), _device.stream(), ts)
    torch.cuda.synchronize()
    return out.cpu().numpy(), int(status.item())


@pytest.mark.parametrize('with_table', [True, False], ids=['phase-table', 'sat'])
def test_static_splat_equals_workspace_splat_twice(gpu, with_table):
    from scopyon_amd.batched import BatchedSimulator
    cases = [([(256, 256)] * 4 + [(96, 160)], [3000, 3000, 3000, 3000, 700], 32),
             ([(128, 128), (64, 64)], [400, 90], 16)]      # second call: smaller input
    for shapes, counts, side in cases:
        print('static splat: shapes {} counts {} tile side {} table {}'.format(
            shapes, counts, side, with_table), flush=True)
        sim = BatchedSimulator(_cfg(shapes[0]), shapes, counts, dt=0.033, seed=sum(counts),
                               tile_side=side if with_table else 32)
        sim.weights()
        want, s_want = _splat(sim, False, with_table)
        got, s_got = _splat(sim, True, with_table)
        assert s_want & ~1 == 0 and s_got == s_want
        assert (want >= 0).all() and want.max() > 0
        assert numpy.array_equal(got, want)


def test_static_splat_rejects_oversized_work(gpu):
    from scopyon_amd import _lib
    from scopyon_amd.batched import BatchedSimulator
    sim = BatchedSimulator(_cfg((64, 64)), [(64, 64)], [20], dt=0.033, seed=1, tile_side=32)
    sim.weights()
    limit = _lib.lib().smx_splat_static_scratch_bytes()
    too_many = limit // 4 + 1          # list entries alone beyond the array
    print('static splat: oversized list capacity {}'.format(too_many), flush=True)
    with pytest.raises(ValueError, match='scratch'):
        _splat(sim, True, True, capacity=too_many)
    # still usable afterwards
    want, _ = _splat(sim, False)
    got, _ = _splat(sim, True)
    assert numpy.array_equal(got, want)
