"""GPU PSF tables vs the oracle: radial profiles and SAT pixel sums."""
import numpy
import pytest
import torch

from oracle import epifm as O

pytestmark = pytest.mark.gpu


def _tables(gpu, wl=5.780000000000062e-07, kind='TRITC', width=None, pl=1.6e-7):
    from scopyon_amd._psf import PSFTables
    return PSFTables('Gaussian' if kind == 'Gaussian' else kind, wl, 1000e-9, 1000e-9,
                     width, pl, gpu)


@pytest.mark.parametrize('depth_nm', [0, 1, 137, 500, 999])
def test_born_wolf_radial_matches_oracle(gpu, depth_nm):
    t = _tables(gpu)
    key = depth_nm
    t.ensure([key])
    r = O.radial_grid(1000e-9)
    ref = O.born_wolf_radial(r, key * 1e-9, 5.780000000000062e-07)
    got = t.radial_last[0].cpu().numpy()
    # float64 with independent j0/sincos: absolute agreement relative to the peak
    assert numpy.abs(got - ref).max() <= 1e-12 * ref.max()


def test_gaussian_radial_matches_oracle(gpu):
    t = _tables(gpu, kind='Gaussian', width=1.0e-7)
    t.ensure([0])
    r = O.radial_grid(1000e-9)
    ref = O.gaussian_radial(r, 1.0e-7)
    got = t.radial_last[0].cpu().numpy()
    numpy.testing.assert_allclose(got, ref, rtol=1e-14, atol=1e-300)


@pytest.mark.parametrize('period,pl', [(66, 6.639004149377593e-08), (160, 1.6e-7)])
def test_sat_pixel_sums_match_bruteforce(gpu, period, pl):
    from scopyon_amd import _lib
    t = _tables(gpu, pl=pl)
    assert t.period == period and t.use_phase == (period == 160)
    t.ensure([0])
    r = O.radial_grid(1000e-9)
    prof = t.radial_last[0].cpu().numpy()
    table = O.radial_to_cartesian(r, prof)          # oracle cartesian table from the GPU profile
    sat = t.sat[t.slot_of[0]].cpu().numpy()
    lib = _lib.lib()
    rng = numpy.random.RandomState(0)
    side = t.ntab + 1
    for _ in range(200):
        a0, a1 = sorted(rng.randint(0, side, 2))
        b0, b1 = sorted(rng.randint(0, side, 2))
        S = lambda a, b: sat[lib.smx_sat_index_host(side, period, int(a), int(b))]
        got = S(a1, b1) - S(a0, b1) - S(a1, b0) + S(a0, b0)
        ref = table[a0:a1, b0:b1].sum()
        assert abs(got - ref) <= 1e-12 * table.sum() + 1e-13 * abs(ref)
