"""The C-ABI library loads on a CPU-only host, exports every symbol declared in
include/*.h, and its host-evaluated helpers are right."""
import ctypes
import os
import re

import numpy
import pytest
from scipy.special import i1e

from scopyon_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for f in os.listdir(os.path.join(ROOT, 'include')):
        if f.endswith('.h'):
            text = open(os.path.join(ROOT, 'include', f)).read()
            names |= set(re.findall(r'\b(smx_\w+)\s*\(', text))
    return names


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert names <= set(_lib.SIGNATURES), names - set(_lib.SIGNATURES)


def test_philox_known_answers():
    """Random123 known-answer vectors for Philox4x32-10."""
    lib = _lib.lib()
    cases = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
             ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
             ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
              (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for ctr, key, want in cases:
        c = (ctypes.c_uint32 * 4)(*ctr)
        k = (ctypes.c_uint32 * 2)(*key)
        o = (ctypes.c_uint32 * 4)()
        lib.smx_philox4x32_10_host(c, k, o)
        assert tuple(o) == want


def test_i1e_against_scipy():
    lib = _lib.lib()
    rng = numpy.random.RandomState(0)
    ys = numpy.concatenate([rng.uniform(0, 1, 500), rng.uniform(0, 8, 500), rng.uniform(8, 50, 500),
                            numpy.exp(rng.uniform(numpy.log(50), numpy.log(1e6), 500)),
                            [0.0, 1e-300, 1e-10, 8.0, numpy.nextafter(8.0, 9.0)]])
    got = numpy.array([lib.smx_i1e_host(float(y)) for y in ys])
    ref = i1e(ys)
    nz = ref != 0
    assert numpy.max(numpy.abs(got[nz] / ref[nz] - 1)) < 2e-15
    assert got[~nz].tolist() == ref[~nz].tolist()


def test_np_remainder_matches_numpy():
    lib = _lib.lib()
    rng = numpy.random.RandomState(1)
    a = numpy.concatenate([rng.normal(0, 10, 1000), [-1e-20, 0.0, -0.0, 5.0, -5.0, 1e300]])
    for b in (1.0, 2.5, 1e-3):
        ref = a % b
        got = numpy.array([lib.smx_np_remainder_host(float(x), b) for x in a])
        assert numpy.array_equal(got, ref)


@pytest.mark.parametrize('side,period', [(2000, 1), (2000, 160), (2000, 66), (2000, 267), (37, 5)])
def test_sat_layout_is_a_bijection(side, period):
    lib = _lib.lib()
    idx = numpy.array([[lib.smx_sat_index_host(side, period, a, b) for b in range(0, side, 7)]
                       for a in range(0, side, 7)])
    assert len(numpy.unique(idx)) == idx.size
    assert idx.max() < lib.smx_psf_sat_elems(side // 2, period) or side % 2


@pytest.mark.parametrize('pl,layout,phase', [
    (160e-9, 160, 160),                  # 16 um / 100x: integer, phase tables
    (16e-6 / 241.0, 66, 1),              # 66.39 cells: layout period only
    (16e-6 / 60.0, 267, 1),              # 266.67 cells rounds up
    (1.2e-9, 1, 1),                      # below two cells: row-major
])
def test_sat_periods(pl, layout, phase):
    """The SAT layout period is the pixel length in cells rounded (corners of
    neighbouring pixels, floor or ceil of it apart, then mostly share a
    residue); phase tables need an exact integer ratio."""
    from scopyon_amd import _psf
    assert _psf.layout_period(pl) == layout
    assert _psf.sat_period(pl) == phase
    if layout > 1:
        # the more frequent of the two edge spacings keeps the residue
        ratio = pl / 1e-9
        edges = numpy.ceil((numpy.arange(40) * ratio + 0.3) - 1e-9).astype(int)
        steps = numpy.diff(edges)
        assert set(steps) <= {int(numpy.floor(ratio)), int(numpy.ceil(ratio))}
        kept = (numpy.diff(edges % layout) == 0).mean()
        assert kept >= 0.5


def test_bad_arguments_raise():
    lib = _lib.lib()
    with pytest.raises(ValueError):
        _lib.call('smx_psf_radial_gaussian', 1, 1e-9, -1.0, None, None)
    msg = lib.smx_last_error().decode()
    assert 'smx_psf_radial_gaussian' in msg


def _pixel_cells(ntab, period, ph):
    """Cell ranges [a, b) of the pixels of phase ph: boundaries 0, ph+1,
    ph+1+P, ... clamped to ntab (the regular grid of _epifm.py:241-263)."""
    edges = [0] + [min(ph + 1 + period * k, ntab) for k in range(ntab // period + 3)]
    cells = [(a, b) for a, b in zip(edges[:-1], edges[1:]) if b > a]
    return cells


@pytest.mark.parametrize('ntab,period', [(1999, 160), (1999, 65), (1999, 2), (101, 10),
                                         (101, 7), (5, 8)])
def test_phase_fold_matches_reflection_of_cell_ranges(ntab, period):
    """The folded phase tables rest on one fact: reflecting the table
    (cell i -> ntab-1-i) maps the pixel cell ranges of phase ph onto those of
    its partner in reverse order.  Checked by brute force on the ranges, with
    the canonical ranks forming a bijection onto the orbits {ph, partner}."""
    lib = _lib.lib()
    rank, refl, npix = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    orbits = {}
    mf = 0
    for ph in range(period):
        _lib.call('smx_phase_fold_host', ntab, period, ph, ctypes.byref(rank), ctypes.byref(refl),
                  ctypes.byref(npix))
        cells = _pixel_cells(ntab, period, ph)
        assert npix.value == len(cells)
        mf = max(mf, npix.value)
        reflected = sorted((ntab - b, ntab - a) for a, b in cells)
        partners = [q for q in range(period) if _pixel_cells(ntab, period, q) == reflected]
        partner = (ntab - 2 - ph) % period
        # several phases share one range set only when the pixel exceeds the table
        assert partner in partners and (len(partners) == 1 or period >= ntab)
        assert refl.value == (partner < ph)
        orbits.setdefault(rank.value, set()).update({ph, partner})
    assert sorted(orbits) == list(range(len(orbits)))
    assert all(len(o) in (1, 2) for o in orbits.values())
    assert sum(len(o) for o in orbits.values()) == period
    nc = len(orbits)
    assert lib.smx_phase_fold_elems(ntab, period) == nc * (nc + 1) // 2 * mf * mf
