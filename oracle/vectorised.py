"""Vectorised-numpy frame of the headline workload: the "best effort CPU"
baseline variant that the benchmark times beside the line-faithful oracle.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``): used by ``bench.py``'s
CPU-baseline leg, never by the package.

Same model as the batched GPU path at 160 nm pixels (an integer number of
1 nm table cells): every emitter's pixel integrals depend only on its two
boundary phases (``phi = ceil(((lo + 1) pl - o) / res)``, the reference's
cell-edge arithmetic of ``_epifm.py:233-275``), so they come from a phase
table K[phi_u][phi_v][du][dv] built once from the summed-area table of the
1 nm cartesian kernel (``epifm.PSFCache``).  The detector is the EMCCD of
``_epifm.py:1436-1469`` drawn with numpy's Generator: n ~ Poisson(E),
S = round(Gamma(n, g)) (>= 1 for n >= 1), Gaussian readout noise, ADC.  Like
the GPU's Philox path this is statistically, not stream-, equivalent to the
reference; emitters whose boundaries fall on a float64 rounding tie
(probability ~1e-12 per boundary) are not special-cased.
"""
import math

import numpy

from .epifm import RES


class PhaseSplat:
    """Phase table of one depth key's cartesian kernel for pixels of P cells."""

    def __init__(self, table, pixel_length):
        self.P = int(round(pixel_length / RES))
        if abs(self.P * RES - pixel_length) > 1e-6 * pixel_length:
            raise ValueError('the vectorised baseline needs an integer pixel/cell ratio')
        self.pl = pixel_length
        self.ntab = table.shape[0]
        S = numpy.zeros((self.ntab + 1, self.ntab + 1))
        S[1:, 1:] = table.cumsum(axis=0).cumsum(axis=1)
        self.M = int(math.ceil(RES * (self.ntab - 1) / pixel_length)) + 2
        P, M, n = self.P, self.M, self.ntab
        m = numpy.arange(M + 1)
        # cell boundary of footprint edge m at phase phi (1..P): 0, then
        # phi + P (m - 1), clamped to the table end
        phi = numpy.arange(1, P + 1)
        B = numpy.where(m[None, :] == 0, 0, phi[:, None] + P * (m[None, :] - 1))
        B = numpy.minimum(B, n)
        b0, b1 = B[:, :-1], B[:, 1:]                     # [phi][du]
        K = (S[b1[:, None, :, None], b1[None, :, None, :]] - S[b0[:, None, :, None], b1[None, :, None, :]]
             - S[b1[:, None, :, None], b0[None, :, None, :]] + S[b0[:, None, :, None], b0[None, :, None, :]])
        K = K * (RES * RES)
        self.K = numpy.where(K > 0.0, K, 0.0)            # [phi_u][phi_v][du][dv]

    def frame(self, shape, u, v, w):
        """Expected photons of emitters at lateral (u, v) relative to the
        focal point with weights w, on a frame of `shape`."""
        nw, nh = shape
        ext_half = (RES * (self.ntab - 1)) * 0.5
        out = numpy.zeros(nw * nh)
        idx, vals = [], []
        lo, ph = [], []
        for x, n_pix in ((u, nw), (v, nh)):
            half = n_pix * self.pl * 0.5
            o = (half + x) - ext_half
            first = numpy.floor(o / self.pl).astype(numpy.int64)
            phi = numpy.ceil(((first + 1) * self.pl - o) / RES).astype(numpy.int64)
            lo.append(first)
            ph.append(numpy.clip(phi, 1, self.P) - 1)
        d = numpy.arange(self.M)
        rows = lo[0][:, None] + d[None, :]                # [e][du]
        cols = lo[1][:, None] + d[None, :]
        ok = ((rows[:, :, None] >= 0) & (rows[:, :, None] < nw) &
              (cols[:, None, :] >= 0) & (cols[:, None, :] < nh))
        k = self.K[ph[0], ph[1]] * w[:, None, None]      # [e][du][dv]
        flat = rows[:, :, None] * nh + cols[:, None, :]
        out += numpy.bincount(flat[ok], weights=k[ok], minlength=nw * nh)
        return out.reshape(nw, nh)


def emccd_adc(photons, p, gen):
    """EMCCD + ADC with a numpy Generator (statistical equivalent of
    _epifm.py:1436-1484)."""
    E = p.QE * (photons + (p.background_mean if p.background_switch else 0.0))
    noise = gen.normal(0.0, p.readout_noise, E.shape) if p.readout_noise > 0 else 0.0
    n = gen.poisson(E)
    S = numpy.zeros(E.shape)
    pos = n > 0
    S[pos] = numpy.maximum(numpy.rint(gen.gamma(n[pos], p.emgain)), 1.0)
    x = numpy.minimum(S + noise, p.adc_fullwell) / p.adc_gain + p.adc_offset
    return numpy.clip(x, 0.0, 2 ** p.adc_bit - 1)
