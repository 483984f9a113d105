"""Pins for the PSF part of the oracle: closed forms, invariants and an
independent high-precision evaluation of the same Riemann sum."""
import math

import numpy
import pytest

from oracle import epifm as O
from oracle import params as P

LAM = 5.780000000000062e-07      # TRITC emission peak on the 1 nm grid


def test_tritc_emission_line_and_normalisation():
    lam, em, _ = P.fluorophore(P.default_config()['fluorophore'])
    assert lam == LAM
    alpha = 2.0 * numpy.pi / lam * 1.4
    # scopyon.yaml:49-51: normalization = 1 / (alpha^2 / pi * (1 nm)^2)
    assert 13564.06122650865 * alpha * alpha / numpy.pi * 1e-18 == pytest.approx(1.0, abs=1e-15)
    assert abs(numpy.sum(em) - 1.0) < 1e-15


def test_born_wolf_closed_form_at_origin():
    r = O.radial_grid(1000e-9)
    psf = O.born_wolf_radial(r, 0.0, LAM)
    alpha = 2.0 * numpy.pi / LAM * 1.4
    # sum_{n=1}^{100} (n/100) * 0.01 = 0.505 -> |I|^2 = 0.255025
    assert psf[0] == pytest.approx(0.255025 * alpha ** 2 / numpy.pi, rel=1e-14)


def test_born_wolf_against_mpmath():
    mpmath = pytest.importorskip('mpmath')
    mpmath.mp.dps = 30
    r = O.radial_grid(1000e-9)
    z = 3.7e-7
    psf = O.born_wolf_radial(r, z, LAM)
    k = 2 * mpmath.pi / LAM
    alpha, gamma = k * 1.4, k * mpmath.mpf(0.49)
    for ri in (0, 1, 123, 400, 777, 999):
        s = mpmath.mpc(0)
        for n in range(1, 101):
            rho = mpmath.mpf(n) / 100
            s += mpmath.besselj(0, alpha * r[ri] * rho) * mpmath.exp(-2j * z * gamma * rho * rho) * rho / 100
        ref = float(abs(s) ** 2 * alpha ** 2 / mpmath.pi)
        assert abs(psf[ri] - ref) <= 1e-12 * psf[0]


def test_born_wolf_depth_symmetry_and_energy():
    r = O.radial_grid(1000e-9)
    a = O.born_wolf_radial(r, 2.5e-7, LAM)
    b = O.born_wolf_radial(r, -2.5e-7, LAM)
    numpy.testing.assert_allclose(a, b, rtol=1e-13, atol=1e-13 * a.max())
    psf0 = O.born_wolf_radial(r, 0.0, LAM)
    # radial integral of the in-focus profile (printed by test/test_epifm.py:30-31)
    tot = numpy.sum(2 * numpy.pi * r * psf0) * 1e-9
    assert 0.9 < tot < 1.0


def test_gaussian_closed_form():
    r = O.radial_grid(1000e-9)
    w = 1.0e-7
    g = O.gaussian_radial(r, w)
    assert g[0] == pytest.approx(1 / (2 * math.pi * w * w), rel=1e-15)
    assert g[100] == pytest.approx(math.exp(-0.5) / (2 * math.pi * w * w), rel=1e-14)


def test_radial_to_cartesian_axes_and_plateau():
    r = O.radial_grid(1000e-9)
    prof = numpy.linspace(3.0, 1.0, r.size)
    tab = O.radial_to_cartesian(r, prof)
    c = r.size - 1
    assert tab.shape == (2 * c + 1, 2 * c + 1)
    assert tab[c, c] == prof[0]
    # exact grid points on the axes reproduce the profile to rounding
    numpy.testing.assert_allclose(tab[c, c:], prof, rtol=1e-13)
    numpy.testing.assert_allclose(tab[c:, c], prof, rtol=1e-13)
    # corners are clamped to the last radial sample (_epifm.py:123)
    assert tab[0, 0] == pytest.approx(prof[-1], rel=1e-13)
    assert numpy.array_equal(tab, tab.T)


def test_vectorised_profile_equals_literal_loop():
    """born_wolf_radial vectorises the reference's per-r map; same bits."""
    from scipy.special import j0
    r = O.radial_grid(1000e-9)[:50]
    z = 1.3e-7
    k = 2.0 * numpy.pi / LAM
    alpha, gamma = k * 1.4, k * numpy.power(1.4 / 2, 2)
    rho = numpy.arange(1, 101) * 0.01
    J0 = numpy.array(list(map(lambda x: j0(x * alpha * rho), r)))
    Y = numpy.exp(-2 * 1.j * z * gamma * rho * rho) * rho * 0.01
    I_sum = (Y * J0).sum(axis=1)
    ref = numpy.array(list(map(lambda x: abs(x) ** 2, I_sum))) * (alpha * alpha / numpy.pi)
    assert numpy.array_equal(O.born_wolf_radial(r, z, LAM), ref)
