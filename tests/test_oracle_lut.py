"""Pins for the LUT-mode oracle (oracle/lut.py): the tabulated pixel
integrals against closed-form totals, exact lattice identities and the
continuous model they approximate; the bilinear splat against the
continuous model, with planted mistakes that fail."""
import math

import numpy
import pytest
from scipy.integrate import quad

from oracle import lut as OL

LAM = 5.780000000000062e-07
PL = 16e-6 / 241.0


@pytest.fixture(scope='module')
def lo():
    return OL.LutOracle(LAM, PL, nz=9)


def test_geometry(lo):
    assert lo.r_max == pytest.approx(999e-9, abs=1e-18)
    # the footprint holds the kernel's square for every phase: c > r_max / pl
    assert lo.c == 16 and lo.m == 33 and lo.c > lo.r_max / PL


def test_blocks_hold_the_whole_kernel(lo):
    """Every phase's footprint covers the square |x|, |y| <= r_max, so each
    block sums to the kernel's integral: the radial profile over the disc plus
    the corner plateau P(r_max) (4 - pi) r_max^2 (_epifm.py:122-125), with
    the square's side as the build grid resolves it (the cells of 1.04 nm
    whose centres lie inside)."""
    cpp = lo.q * lo.ns
    h = PL / cpp
    centres = -(lo.c + 1) * PL + (numpy.arange((lo.m + 1) * cpp) + 0.5) * h
    side = numpy.count_nonzero(numpy.abs(centres) <= lo.r_max) * h
    assert abs(side - 2 * lo.r_max) < 2 * h
    for s in (0, 4, 8):
        prof = lo.profile(s)
        r = numpy.arange(lo.nb) * (lo.r_max / (lo.nb - 1))
        disc, _ = quad(lambda x: 2 * math.pi * x * numpy.interp(x, r, prof), 0, lo.r_max,
                       limit=400, points=list(r[::64]))
        total = disc + prof[-1] * (side * side - math.pi * lo.r_max ** 2)
        for qu, qv in ((0, 0), (7, 30), (32, 1)):
            assert lo.block(s, qu, qv).sum() == pytest.approx(total, rel=2e-6), (s, qu, qv)


def test_phase_lattice_identities(lo):
    """Phase q of pixel du is phase 0 of pixel du - 1 (same cells); the
    kernel's transpose and mirror symmetries carry over to the blocks."""
    a = lo.block(2, 32, 5)
    b = lo.block(2, 0, 5)
    assert numpy.array_equal(a[1:, :], b[:-1, :])
    assert not a[0, :].any() and not b[-1, :].any()
    t = lo.block(2, 5, 11)
    numpy.testing.assert_allclose(lo.block(2, 11, 5), t.T, rtol=1e-12, atol=1e-13 * t.max())
    # s -> -s maps pixel du at phase qu onto pixel m - 1 - du at phase q - qu
    mirror = lo.block(2, 32 - 5, 11)[::-1, :]
    numpy.testing.assert_allclose(mirror, t, rtol=1e-12, atol=1e-13 * t.max())


def test_blocks_equal_continuous_integral_at_grid_phases(lo):
    """An emitter exactly on a tabulated phase: the block (midpoint rule on
    64 x 64 cells per pixel) equals the continuous pixel integral (48 x 48
    nodes) of the same kernel to quadrature accuracy."""
    x_rel, y_rel = 20 + 7 / 32, 20 + 19 / 32
    i0, j0, cont = OL.pixel_integral_exact(lo, 0.0, x_rel, y_rel)
    blk = lo.block(0, 7, 19)
    assert (i0, j0) == (20 - lo.c, 20 - lo.c)
    assert numpy.abs(blk - cont).max() <= 2e-5 * cont.max()


def _frame(lo, u, v, w, depth=None, swap=False):
    shape = (40, 40)
    depth = numpy.zeros(len(u)) if depth is None else depth
    if swap:
        real = OL.emitter_phase
        try:
            OL.emitter_phase = lambda x, half, pl, q: (lambda r: (r[0], q - 1 - r[1], r[2]))(
                real(x, half, pl, q))
            return OL.splat_lut(lo, shape, depth, u, v, w, 0.0, 0.0, 0.0)
        finally:
            OL.emitter_phase = real
    return OL.splat_lut(lo, shape, depth, u, v, w, 0.0, 0.0, 0.0)


def test_bilinear_splat_tracks_continuous_model(lo):
    """Emitters between tabulated phases (in focus, where the profile is
    sharpest): within 2e-3 of the frame maximum of the continuous model; a
    phase taken from the wrong end of the pixel fails."""
    rng = numpy.random.RandomState(5)
    u = rng.uniform(-8 * PL, 8 * PL, 3)
    v = rng.uniform(-8 * PL, 8 * PL, 3)
    w = numpy.array([1.0, 0.7, 1.3])
    got = _frame(lo, u, v, w)
    cont = numpy.zeros((40, 40))
    half = 40 * PL * 0.5
    for i in range(3):
        i0, j0, blk = OL.pixel_integral_exact(lo, 0.0, (half + u[i]) / PL, (half + v[i]) / PL)
        a0, b0 = max(i0, 0), max(j0, 0)
        a1, b1 = min(i0 + lo.m, 40), min(j0 + lo.m, 40)
        cont[a0:a1, b0:b1] += w[i] * blk[a0 - i0:a1 - i0, b0 - j0:b1 - j0]
    assert numpy.abs(got - cont).max() <= 2e-3 * cont.max()
    bad = _frame(lo, u, v, w, swap=True)
    assert numpy.abs(bad - cont).max() > 1e-2 * cont.max()


def test_slice_of_depth():
    cut = 1000e-9
    assert OL.slice_of_depth(0.0, 0.0, cut, 256) == 0
    assert OL.slice_of_depth(-1000e-9, 0.0, cut, 256) == 255
    assert OL.slice_of_depth(5e-6, 0.0, cut, 256) == 255
    dz = cut / 255
    # 1 nm key first (int(|d| / 1 nm)), then the nearest slice
    assert OL.slice_of_depth(1.9 * dz, 0.0, cut, 256) == 2
    assert OL.slice_of_depth(0.4 * dz, 0.0, cut, 256) == 0
