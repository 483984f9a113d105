"""float64 restatement of the LUT-mode PSF (radial x depth lookup table) and
of the continuous pixel integral it approximates.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

LUT mode is this package's deliberate deviation from the reference's 1 nm
depth keys and 1 nm cell sums (``_epifm.py:76-126, 215-282``): the same
cartesian kernel -- the Born-Wolf radial profile (``_epifm.py:180-213``,
``epifm.born_wolf_radial``) clamped at r_max on the square |x|, |y| <= r_max
(``_epifm.py:98-126``) -- integrated over the exact pixel square and
tabulated per depth slice at (q + 1)^2 sub-pixel phases.  Three levels:

* ``radial_lut``, ``fine_sat``, ``phase_block``: the tables themselves
  (midpoint rule on cells of pl / (q ns), summed-area table in float64, one
  four-corner difference per entry), for checking the device tables;
* ``splat_lut``: frames from those tables with the device's phase/fraction
  arithmetic (float32 bilinear taps), for checking the device splat;
* ``pixel_integral_exact``: the continuous integral at the emitter's exact
  position and depth (no slice, phase or cell rounding), the model the LUT
  mode approximates, for its error budget.

The parity path (1 nm tables) is ``epifm.molecule_plane``.
"""
import math

import numpy

from .epifm import RES, born_wolf_radial


def lut_geometry(pixel_length, radial_cutoff=1000e-9):
    """(r_max, c, m): the reference's largest radial sample (_epifm.py:91),
    the footprint pixel of the emitter and the footprint side."""
    nr = len(numpy.arange(0.0, radial_cutoff, RES, dtype=float))
    r_max = (nr - 1) * RES
    c = int(math.floor(r_max / pixel_length)) + 1
    return r_max, c, 2 * c + 1


def slice_depths(depth_cutoff, nz):
    return numpy.linspace(0.0, depth_cutoff, nz) if nz > 1 else numpy.zeros(1)


def slice_of_depth(depth, focal, depth_cutoff, nz):
    """Depth slice of an emitter: its 1 nm depth key (_epifm.py:76-85; the
    cutoff key above cutoff + 1 nm), then the nearest slice."""
    if nz == 1:
        return 0
    d = abs(depth - focal)
    if not d < depth_cutoff + RES:
        return nz - 1
    key = int(d / RES)
    dz = depth_cutoff / (nz - 1)
    return int(min(nz - 1, round(key * RES / dz)))


def radial_lut(z, wave_length, r_max, nb):
    """Born-Wolf profile on nb bins over [0, r_max] at depth z."""
    r = numpy.arange(nb) * (r_max / (nb - 1))
    return born_wolf_radial(r, z, wave_length)


def kernel(prof, r_max, x, y):
    """Cartesian kernel: linear interpolation of the profile at
    min(sqrt(x^2 + y^2), r_max) on the square |x|, |y| <= r_max, 0 outside."""
    nb = prof.size
    dr = r_max / (nb - 1)
    R = numpy.minimum(numpy.sqrt(x * x + y * y), r_max)
    t = R / dr
    k = numpy.minimum(t.astype(numpy.int64), nb - 2)
    f = t - k
    v = prof[k] + f * (prof[k + 1] - prof[k])
    return numpy.where((numpy.abs(x) <= r_max) & (numpy.abs(y) <= r_max), v, 0.0)


def fine_sat(prof, r_max, pixel_length, q, ns, c, m):
    """Summed-area table of the kernel at the cell centres of the build grid:
    cells of pl / (q ns) covering [-(c + 1) pl, (c + 1) pl) per axis."""
    cpp = q * ns
    side = (m + 1) * cpp
    h = pixel_length / cpp
    centres = -(c + 1) * pixel_length + (numpy.arange(side) + 0.5) * h
    vals = kernel(prof, r_max, centres[None, :], centres[:, None])   # [row = x][col = y]
    S = numpy.zeros((side + 1, side + 1))
    S[1:, 1:] = vals.cumsum(axis=1).cumsum(axis=0)
    return S, h


def phase_block(S, h, q, ns, m, qu, qv):
    """K[qu][qv][du][dv]: pixel du spans [du - c - qu/q, du - c + 1 - qu/q)
    pixels from the emitter."""
    cpp = q * ns
    du = numpy.arange(m)
    u0 = (du + 1) * cpp - qu * ns
    v0 = (du + 1) * cpp - qv * ns
    U0, V0 = u0[:, None], v0[None, :]
    U1, V1 = U0 + cpp, V0 + cpp
    return (((S[U1, V1] - S[U0, V1]) - S[U1, V0]) + S[U0, V0]) * (h * h)


class LutOracle:
    """Tables on demand (per slice) for one optical configuration."""

    def __init__(self, wave_length, pixel_length, depth_cutoff=1000e-9, nz=256, nb=4096, q=32,
                 ns=2, radial_cutoff=1000e-9):
        self.wave_length, self.pl = wave_length, pixel_length
        self.depth_cutoff, self.nz, self.nb, self.q, self.ns = depth_cutoff, nz, nb, q, ns
        self.r_max, self.c, self.m = lut_geometry(pixel_length, radial_cutoff)
        self.depths = slice_depths(depth_cutoff, nz)
        self._sat = {}

    def profile(self, s):
        return radial_lut(self.depths[s], self.wave_length, self.r_max, self.nb)

    def sat(self, s):
        if s not in self._sat:
            self._sat[s] = fine_sat(self.profile(s), self.r_max, self.pl, self.q, self.ns, self.c,
                                    self.m)
        return self._sat[s]

    def block(self, s, qu, qv):
        S, h = self.sat(s)
        return phase_block(S, h, self.q, self.ns, self.m, qu, qv)


def _f32(x):
    return numpy.float32(x)


def emitter_phase(x, half, pixel_length, q):
    """(first pixel of the emitter's pixel, phase, 16-bit fraction) from the
    device's float64 arithmetic: t = (half + x) / pl * q."""
    t = (half + x) / pixel_length * q
    qi = math.floor(t)
    pix = qi // q
    return pix, qi - pix * q, int((t - qi) * 65535.0 + 0.5)


def splat_lut(lo, shape, depth, u, v, w, focal, p0u, p0v):
    """Frame (nw, nh) of emitters (depth, u, v, weight w) from the LUT
    tables: per emitter its slice, phase and fraction, the float32 bilinear
    combination of the four tabulated phases (the device's operation order),
    added to the float64 frame in emitter order."""
    nw, nh = shape
    out = numpy.zeros((nw, nh))
    half_w, half_h = nw * lo.pl * 0.5, nh * lo.pl * 0.5
    inv = _f32(1.0 / 65535.0)
    for i in range(len(u)):
        if not w[i] > 0.0:
            continue
        s = slice_of_depth(depth[i], focal, lo.depth_cutoff, lo.nz)
        pu, qu, fu16 = emitter_phase(u[i] - p0u, half_w, lo.pl, lo.q)
        pv, qv, fv16 = emitter_phase(v[i] - p0v, half_h, lo.pl, lo.q)
        fu, fv, wf = _f32(fu16) * inv, _f32(fv16) * inv, _f32(w[i])
        one = _f32(1.0)
        c = [((one - fu) * (one - fv)) * wf, ((one - fu) * fv) * wf, (fu * (one - fv)) * wf,
             (fu * fv) * wf]
        K = [lo.block(s, qu + a, qv + b).astype(numpy.float32) for a, b in
             ((0, 0), (0, 1), (1, 0), (1, 1))]
        val = ((c[0] * K[0] + c[1] * K[1]) + c[2] * K[2]) + c[3] * K[3]
        i0, j0 = pu - lo.c, pv - lo.c
        a0, b0 = max(i0, 0), max(j0, 0)
        a1, b1 = min(i0 + lo.m, nw), min(j0 + lo.m, nh)
        if a1 <= a0 or b1 <= b0:
            continue
        out[a0:a1, b0:b1] += val[a0 - i0:a1 - i0, b0 - j0:b1 - j0].astype(numpy.float64)
    return out


def pixel_integral_exact(lo, depth_rel, x_rel, y_rel, n=48):
    """Continuous model: the kernel at depth |depth_rel| (its own Born-Wolf
    profile, no slice) integrated over each footprint pixel by an n x n
    midpoint rule, emitter at (x_rel, y_rel) pixels from the frame edge.
    Returns (first pixel u, first pixel v, m x m block)."""
    prof = radial_lut(abs(depth_rel), lo.wave_length, lo.r_max, lo.nb)
    pu, pv = int(math.floor(x_rel)), int(math.floor(y_rel))
    i0, j0 = pu - lo.c, pv - lo.c
    sub = (numpy.arange(n) + 0.5) / n
    out = numpy.zeros((lo.m, lo.m))
    for a in range(lo.m):
        X = (i0 + a + sub - x_rel) * lo.pl
        for b in range(lo.m):
            Y = (j0 + b + sub - y_rel) * lo.pl
            out[a, b] = kernel(prof, lo.r_max, X[:, None], Y[None, :]).sum() * (lo.pl / n) ** 2
    return i0, j0, out
