"""Pins of the float64 Gaussian-model oracle (oracle/gauss.py): pixel factors
against numerical quadrature, their telescoping sum, and the analytic
gradients against central finite differences."""
import numpy
import pytest
from scipy.integrate import quad
from scipy.special import erf

from oracle import gauss as G


@pytest.mark.parametrize('x,sigma', [(3.3, 1.2), (0.5, 0.7), (10.0, 2.0), (7.91, 1.0)])
def test_g_pixel_is_the_pixel_integral_of_the_normal_density(x, sigma):
    i = numpy.arange(int(x) - 9, int(x) + 10)
    got = G.g_pixel(i, x, sigma)
    want = numpy.array([quad(lambda t: G.normal_pdf(t - x, sigma), a, a + 1.0,
                             epsabs=0.0, epsrel=1e-13)[0] for a in i])
    numpy.testing.assert_allclose(got, want, rtol=1e-9, atol=0.0)


def test_g_pixel_window_sum_telescopes():
    x, sigma, lo, W = 5.37, 1.3, 1, 9
    got = G.g_pixel(lo + numpy.arange(W), x, sigma).sum()
    s = numpy.sqrt(2.0) * sigma
    want = 0.5 * (erf((lo + W - x) / s) - erf((lo - x) / s))
    assert abs(got - want) < 1e-15
    wide = G.g_pixel(numpy.arange(-40, 50), x, sigma).sum()
    assert abs(wide - 1.0) < 1e-15


def _case(seed=0):
    rng = numpy.random.RandomState(seed)
    shapes = [(24, 20), (18, 26)]
    n = 7
    x = rng.uniform(-1.0, 23.0, n)
    y = rng.uniform(-1.0, 20.0, n)
    x[0], y[0] = 0.3, 19.6          # window crosses two frame edges
    z = rng.uniform(-0.5, 0.5, n)
    I = rng.uniform(200.0, 800.0, n)
    fid = numpy.array([0, 0, 0, 1, 1, 1, 0])
    kw = dict(sigma=1.2, inv_depth=1.0 / 0.377, background=5.0, radius=5)
    truth = G.render(shapes, x + 0.2, y - 0.1, z, I, fid, **kw)
    targets = [rng.poisson(m).astype(float) for m in truth]
    return shapes, x, y, z, I, fid, kw, targets


def _fd(shapes, x, y, z, I, fid, kw, targets, which, e, h):
    args = dict(x=x.copy(), y=y.copy(), z=z.copy(), I=I.copy())
    kw = dict(kw)
    def L(delta):
        a = {k: v.copy() for k, v in args.items()}
        k2 = dict(kw)
        if which == 'sigma':
            k2['sigma'] = kw['sigma'] + delta
        else:
            a[which][e] += delta
        return G.loss(shapes, a['x'], a['y'], a['z'], a['I'], fid, targets=targets, **k2)
    return (L(h) - L(-h)) / (2 * h)


def test_analytic_gradients_match_central_differences():
    shapes, x, y, z, I, fid, kw, targets = _case()
    gx, gy, gz, gI, gsig = G.gradients(shapes, x, y, z, I, fid, targets=targets, **kw)
    for e in range(len(x)):
        for which, got, h in (('x', gx[e], 1e-4), ('y', gy[e], 1e-4), ('z', gz[e], 1e-5),
                              ('I', gI[e], 1e-3)):
            fd = _fd(shapes, x, y, z, I, fid, kw, targets, which, e, h)
            assert abs(got - fd) <= 1e-6 * max(1.0, abs(fd)), (which, e, got, fd)
    fd = _fd(shapes, x, y, z, I, fid, kw, targets, 'sigma', 0, 1e-5)
    assert abs(gsig - fd) <= 1e-6 * max(1.0, abs(fd)), (gsig, fd)


def test_planted_density_mistake_is_caught(monkeypatch):
    """A 1% error in the normal density (used only by the analytic
    derivatives) must fail the finite-difference comparison."""
    shapes, x, y, z, I, fid, kw, targets = _case(1)
    real = G.normal_pdf
    monkeypatch.setattr(G, 'normal_pdf', lambda t, s: 1.01 * real(t, s))
    gx, _, _, _, _ = G.gradients(shapes, x, y, z, I, fid, targets=targets, **kw)
    monkeypatch.setattr(G, 'normal_pdf', real)
    bad = 0
    for e in range(len(x)):
        fd = _fd(shapes, x, y, z, I, fid, kw, targets, 'x', e, 1e-4)
        bad += abs(gx[e] - fd) > 1e-6 * max(1.0, abs(fd))
    assert bad >= len(x) - 1
