"""GPU EMCCD + ADC vs the oracle with the same expectation and random numbers."""
import numpy
import pytest
import torch

from oracle import epifm as O
from conftest import oracle_params

pytestmark = pytest.mark.gpu


def _det_params(p):
    from scopyon_amd import _lib
    d = _lib.DetectorParamsT()
    d.qe, d.background, d.background_switch = p.QE, p.background_mean, 1
    d.emgain, d.readout_noise = p.emgain, p.readout_noise
    d.fullwell, d.adc_max = p.adc_fullwell, 2 ** p.adc_bit - 1
    d.offset, d.gain = float(p.adc_offset.flat[0]), float(p.adc_gain.flat[0])
    return d


def _gpu_parity(gpu, E, noise, uniforms, det, gain_map=None, offset_map=None):
    """Run the parity EMCCD kernel on a given expectation array."""
    from scopyon_amd import _lib, _device
    n = E.size
    dE = _device.upload(E.ravel(), device=gpu)
    flag = (E.ravel() > 0).astype(numpy.int32)
    uidx = numpy.concatenate([[0], numpy.cumsum(flag)[:-1]]).astype(numpy.int32)
    d_uidx = _device.upload(uidx, dtype=torch.int32, device=gpu)
    d_u = _device.upload(uniforms if len(uniforms) else numpy.zeros(1), device=gpu)
    d_noise = None if noise is None else _device.upload(noise.ravel(), device=gpu)
    adc = torch.empty(n, dtype=torch.float64, device=gpu)
    sig = torch.empty(n, dtype=torch.float64, device=gpu)
    status = torch.zeros(1, dtype=torch.int32, device=gpu)
    gm = None if gain_map is None else _device.upload(gain_map.ravel(), device=gpu)
    om = None if offset_map is None else _device.upload(offset_map.ravel(), device=gpu)
    _lib.call('smx_emccd_parity', n, dE.data_ptr(), _device.ptr(d_noise), d_u.data_ptr(),
              d_uidx.data_ptr(), _device.ptr(gm), _device.ptr(om), det, -1.0, None,
              adc.data_ptr(), sig.data_ptr(), status.data_ptr(), _device.stream())
    return adc.cpu().numpy().reshape(E.shape), sig.cpu().numpy().reshape(E.shape), int(status.item())


def _oracle(p, E, noise, uniforms, gain=None, offset=None):
    S = O.emccd_signal_predrawn(E, p.emgain, uniforms)
    adc = O.adc_counts(S + noise, p.adc_fullwell, p.adc_gain if gain is None else gain,
                       p.adc_offset if offset is None else offset, p.adc_bit)
    return adc, S


@pytest.mark.parametrize('seed,scale', [(0, 0.0), (1, 1.0), (2, 10.0), (3, 100.0), (4, 600.0)])
def test_emccd_parity_bit_exact(gpu, seed, scale):
    rng = numpy.random.RandomState(seed)
    shape = (24, 20)
    p = oracle_params({'detector': {'image_size': list(shape)}})
    E = 0.0092 + scale * rng.exponential(1.0, size=shape) * (rng.uniform(size=shape) < 0.5)
    E = numpy.minimum(E, 700.0)   # above ~709.8 the reference's pmf overflows (tested below)
    noise = rng.normal(0, p.readout_noise, shape)
    u = rng.random_sample(int((E > 0).sum()))
    ref_adc, ref_S = _oracle(p, E, noise, u)
    adc, S, status = _gpu_parity(gpu, E, noise, u, _det_params(p))
    assert status == 0
    assert numpy.array_equal(S, ref_S)
    assert numpy.array_equal(adc, ref_adc)


def test_emccd_parity_zero_and_edge_uniforms(gpu):
    p = oracle_params({'detector': {'image_size': [4, 4]}})
    E = numpy.array([[0.0, 0.0092, 0.5, 3.0], [12.0, 0.0, 1e-6, 50.0],
                     [0.0092, 0.0092, 0.0092, 0.0092], [2.0, 2.0, 2.0, 2.0]])
    u = numpy.array([0.0, 0.999999999999, 0.5, 0.25, 0.75, 1e-300, 0.9,
                     0.0, 0.99084, 0.9909, 0.5, 0.1, 0.2, 0.3])[: int((E > 0).sum())]
    noise = numpy.zeros_like(E)
    ref_adc, ref_S = _oracle(p, E, noise, u)
    adc, S, status = _gpu_parity(gpu, E, noise, u, _det_params(p))
    assert status == 0
    assert numpy.array_equal(S, ref_S)
    assert numpy.array_equal(adc, ref_adc)


def test_emccd_parity_common_total_and_long_supports(gpu):
    """Pixels sharing the background-only expectation search the pmf's chunk
    sums computed once (smx_emccd_common_total); with EM gain 1000 the
    supports of E ~ 60-150 exceed the chunk store and take the unchunked
    passes.  Both against the oracle's RandomState.choice draws."""
    from scopyon_amd import _lib, _device
    rng = numpy.random.RandomState(12)
    for emgain, E_common, others in ((300, 0.0092, [0.3, 4.0, 37.0, 250.0]),
                                     (1000, 0.92, [0.2, 3.0, 64.0, 150.0])):
        print('parity common total: emgain {}'.format(emgain), flush=True)
        p = oracle_params({'detector': {'image_size': [8, 8], 'emgain': emgain}})
        E = numpy.full((8, 8), E_common)
        E.flat[[3, 17, 40, 63]] = others
        noise = rng.normal(0, p.readout_noise, E.shape)
        u = rng.random_sample(E.size)
        ref_adc, ref_S = _oracle(p, E, noise, u)
        det = _det_params(p)
        n = E.size
        T = torch.empty(int(_lib.lib().smx_emccd_common_elems()), dtype=torch.float64,
                        device=gpu)
        _lib.call('smx_emccd_common_total', E_common, emgain, T.data_ptr(), _device.stream())
        dE = _device.upload(E.ravel(), device=gpu)
        d_uidx = _device.upload(numpy.arange(n), dtype=torch.int32, device=gpu)
        d_u = _device.upload(u, device=gpu)
        d_noise = _device.upload(noise.ravel(), device=gpu)
        adc = torch.empty(n, dtype=torch.float64, device=gpu)
        sig = torch.empty(n, dtype=torch.float64, device=gpu)
        status = torch.zeros(1, dtype=torch.int32, device=gpu)
        _lib.call('smx_emccd_parity', n, dE.data_ptr(), d_noise.data_ptr(), d_u.data_ptr(),
                  d_uidx.data_ptr(), None, None, det, E_common, T.data_ptr(), adc.data_ptr(),
                  sig.data_ptr(), status.data_ptr(), _device.stream())
        assert int(status.item()) == 0
        assert numpy.array_equal(sig.cpu().numpy().reshape(E.shape), ref_S)
        assert numpy.array_equal(adc.cpu().numpy().reshape(E.shape), ref_adc)


def test_emccd_parity_fpn_maps(gpu):
    rng = numpy.random.RandomState(3)
    p = oracle_params({'detector': {'image_size': [16, 16]},
                       'analog_to_digital_converter': {'type': 'pixel', 'count': 30.0}},
                      rng=numpy.random.RandomState(8))
    E = rng.exponential(3.0, (16, 16))
    noise = rng.normal(0, 100, (16, 16))
    u = rng.random_sample(256)
    ref_adc, _ = _oracle(p, E, noise, u)
    adc, _, status = _gpu_parity(gpu, E, noise, u, _det_params(p), p.adc_gain, p.adc_offset)
    assert status == 0
    assert numpy.array_equal(adc, ref_adc)


def test_emccd_parity_overflow_is_reported(gpu):
    # exp(-X + Y) overflows near X = E for E > ~709.8: numpy.random.choice raises there
    p = oracle_params({'detector': {'image_size': [1, 2]}})
    E = numpy.array([[5.0, 800.0]])
    with pytest.raises(ValueError), numpy.errstate(all='ignore'):
        O.emccd_signal(E, p.emgain, numpy.random.RandomState(0))
    _, _, status = _gpu_parity(gpu, E, None, numpy.array([0.5, 0.5]), _det_params(p))
    assert status & 4


def test_emccd_philox_moments(gpu):
    """Philox EMCCD: mean g E and variance 2 g^2 E (excess noise factor 2),
    atom P(S=0) = exp(-E), readout noise N(0, sigma^2)."""
    from scopyon_amd import _lib, _device
    p = oracle_params({'detector': {'image_size': [512, 512]}})
    det = _det_params(p)
    det.readout_noise = 0.0
    det.background_switch = 0
    n = 512 * 512
    for E_val in (0.05, 2.0, 30.0):
        photons = torch.full((n,), E_val / p.QE, dtype=torch.float64, device=gpu)
        adc = torch.empty(n, dtype=torch.float64, device=gpu)
        E = torch.empty(n, dtype=torch.float64, device=gpu)
        det.gain, det.offset, det.fullwell, det.adc_max = 1.0, 0.0, 1e300, 1e300
        _lib.call('smx_emccd_philox', n, photons.data_ptr(), None, None, det, 1234, 7, 0, None,
                  E.data_ptr(), adc.data_ptr(), 0, _device.stream())
        S = adc.cpu().numpy()
        g = p.emgain
        Ev = E.cpu().numpy()[0]
        assert abs(S.mean() - g * Ev) < 5 * numpy.sqrt(2 * g * g * Ev / n)
        assert abs(S.var() / (2 * g * g * Ev) - 1) < 0.08
        assert abs((S == 0).mean() - numpy.exp(-Ev)) < 5 * numpy.sqrt(numpy.exp(-Ev) / n) + 1e-4
        assert (S == numpy.floor(S)).all()


def test_emccd_philox_readout_noise_is_gaussian(gpu):
    """Readout noise of the Philox path (Box-Muller on the hardware log2 and
    cos): N(0, sigma^2) by mean, variance and a Kolmogorov-Smirnov test."""
    from scipy import stats
    from scopyon_amd import _lib, _device
    p = oracle_params({'detector': {'image_size': [512, 512]}})
    det = _det_params(p)
    det.background_switch = 0
    det.readout_noise = 100.0
    det.gain, det.offset, det.fullwell, det.adc_max = 1.0, 1000.0, 1e300, 1e300
    n = 512 * 512
    photons = torch.zeros(n, dtype=torch.float64, device=gpu)
    adc = torch.empty(n, dtype=torch.float32, device=gpu)
    _lib.call('smx_emccd_philox', n, photons.data_ptr(), None, None, det, 99, 3, 0, None, None,
              adc.data_ptr(), 1, _device.stream())
    x = (adc.cpu().numpy().astype(float) - 1000.0) / 100.0
    assert abs(x.mean()) < 5 / numpy.sqrt(n)
    assert abs(x.var() - 1.0) < 5 * numpy.sqrt(2.0 / n)
    assert stats.kstest(x, 'norm').pvalue > 1e-3


@pytest.mark.parametrize('E_val', [2.0, 30.0])
def test_emccd_philox_distribution(gpu, E_val):
    """Philox EMCCD register against the law it samples: k ~ Poisson(E), then
    S = round(Gamma(k, g)) redrawn below 0.5 (S = 0 when k = 0), so
    P(S <= s) = exp(-E) + sum_k pois(k) (F_k(s + 1/2) - F_k(1/2)) / (1 - F_k(1/2))
    with F_k the Gamma(k, scale g) cdf.  Pearson chi-square over the atom at 0
    (merged into the first bin when it expects < 20 pixels) and 30 bins of
    equal probability (262,144 pixels, fixed seed) must not reject at 1e-4; a
    1% error in the gain is rejected (checked with numpy's samplers)."""
    from scipy import stats
    from scopyon_amd import _lib, _device
    p = oracle_params({'detector': {'image_size': [512, 512]}})
    det = _det_params(p)
    det.readout_noise = 0.0
    det.background_switch = 0
    det.gain, det.offset, det.fullwell, det.adc_max = 1.0, 0.0, 1e300, 1e300
    n = 512 * 512
    photons = torch.full((n,), E_val / p.QE, dtype=torch.float64, device=gpu)
    adc = torch.empty(n, dtype=torch.float64, device=gpu)
    E = torch.empty(n, dtype=torch.float64, device=gpu)
    _lib.call('smx_emccd_philox', n, photons.data_ptr(), None, None, det, 2024, 11, 0, None,
              E.data_ptr(), adc.data_ptr(), 0, _device.stream())
    S = adc.cpu().numpy()
    Ev = float(E.cpu().numpy()[0])
    g = float(p.emgain)
    assert (S == numpy.floor(S)).all() and (S >= 0).all()
    ks = numpy.arange(1, int(Ev + 12 * numpy.sqrt(Ev) + 20))
    wk = stats.poisson.pmf(ks, Ev)
    s_max = int(g * (ks[-1] + 10 * numpy.sqrt(ks[-1])))
    sv = numpy.arange(0, s_max + 1, dtype=float)
    F = stats.gamma.cdf(sv[None, :] + 0.5, ks[:, None], scale=g)
    F0 = stats.gamma.cdf(0.5, ks, scale=g)[:, None]
    cdf = numpy.exp(-Ev) + (wk[:, None] * (F - F0) / (1.0 - F0)).sum(axis=0)   # P(S <= s)
    p0 = numpy.exp(-Ev)
    # the atom at 0 is a bin of its own when it expects >= 20 pixels
    atom = p0 * n >= 20
    lo = 0 if atom else -1
    targets = p0 + (1.0 - p0) * numpy.arange(1, 30) / 30.0
    uppers = numpy.unique(numpy.searchsorted(cdf, targets))        # bins (prev, upper]
    edges = numpy.concatenate([[lo], uppers, [s_max + 10 ** 9]])
    obs = [((S > a) & (S <= b)).sum() for a, b in zip(edges[:-1], edges[1:])]
    probs = numpy.diff(numpy.concatenate([[p0 if atom else 0.0], cdf[uppers], [1.0]]))
    if atom:
        obs, probs = [(S == 0).sum()] + obs, numpy.concatenate([[p0], probs])
    obs = numpy.array(obs)
    exp = probs * n
    assert obs.sum() == n and exp.min() > 20
    chi2 = float(((obs - exp) ** 2 / exp).sum())
    pval = stats.chi2.sf(chi2, len(exp) - 1)
    assert pval > 1e-4, (E_val, chi2, len(exp), pval)


@pytest.mark.parametrize('lam', [0.5, 5.0, 11.9, 12.0, 30.0, 150.0, 2000.0])
def test_philox_poisson_sampler_distribution(gpu, lam):
    """The device Poisson sampler (inversion below 12, PTRS from 12 up, shared
    by the EMCCD and CCD/CMOS paths) against the Poisson pmf: a CCD readout
    with no noise, unit gain and no offset returns the count itself.  Pearson
    chi-square over bins of expected count >= 20 (262,144 draws, fixed seed)
    must not reject at 1e-4; mean and variance within 5 standard errors."""
    from scipy import stats
    from scopyon_amd import _lib, _device
    det = _lib.DetectorParamsT()
    det.qe, det.background, det.background_switch, det.emgain = 1.0, 0.0, 0, 1
    det.readout_noise, det.fullwell, det.adc_max = 0.0, 1e300, 1e300
    det.offset, det.gain = 0.0, 1.0
    n = 512 * 512
    photons = torch.full((n,), lam, dtype=torch.float64, device=gpu)
    adc = torch.empty(n, dtype=torch.float64, device=gpu)
    _lib.call('smx_photon_detector_philox', n, photons.data_ptr(), None, None, det, None, None, 0,
              4321, 5, 0, None, None, adc.data_ptr(), 0, _device.stream())
    k = adc.cpu().numpy()
    assert (k == numpy.floor(k)).all() and (k >= 0).all()
    assert abs(k.mean() - lam) < 5 * numpy.sqrt(lam / n)
    assert abs(k.var() / lam - 1.0) < 5 * numpy.sqrt(2.0 / n) + 1e-3
    # bins [lo, hi) merged from both tails until each expects >= 20 draws
    ks = numpy.arange(0, int(lam + 12 * numpy.sqrt(lam) + 20))
    pmf = stats.poisson.pmf(ks, lam)
    edges, acc = [0], 0.0
    for i, p in enumerate(pmf):
        acc += p
        if acc * n >= 20 and (1.0 - pmf[:i + 1].sum()) * n >= 20:
            edges.append(i + 1)
            acc = 0.0
    edges.append(len(ks) + 10 ** 9)
    obs = numpy.histogram(k, bins=numpy.array(edges, dtype=float) - 0.5)[0]
    cdf = stats.poisson.cdf(numpy.array(edges[1:-1]) - 1, lam)
    exp = numpy.diff(numpy.concatenate([[0.0], cdf, [1.0]])) * n
    chi2 = float(((obs - exp) ** 2 / exp).sum())
    pval = stats.chi2.sf(chi2, len(exp) - 1)
    assert pval > 1e-4, (lam, chi2, len(exp), pval)
