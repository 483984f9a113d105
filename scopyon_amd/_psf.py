"""Device-resident PSF tables: radial profiles and summed-area tables (SATs),
one per 1 nm depth key, built on first use like the reference's per-key cache
(src/scopyon/_epifm.py:76-96) but kept in HBM across frames and calls.

A SAT holds the (2 n - 1)^2 cartesian table of ``radial_to_cartesian``
(_epifm.py:98-126) as corner prefix sums; the splat kernel integrates any
pixel rectangle from four of its entries.  One table is 32 MB at the default
1 nm grid, so even all 1002 depth keys (32 GB) fit in one MI355X's HBM.
"""
import math

import numpy
import torch

from . import _device, _lib

RES = 1.0e-9   # _epifm.py:59-60


def depth_keys(depth, focal_depth, depth_cutoff):
    """Vectorised ``PointSpreadingFunction.get`` key (_epifm.py:76-85); -1 = cutoff."""
    ad = numpy.abs(numpy.asarray(depth, dtype=numpy.float64) - focal_depth)
    keys = numpy.full(ad.shape, -1, dtype=numpy.int64)
    inside = ad < depth_cutoff + RES
    keys[inside] = (ad[inside] / RES).astype(numpy.int64)
    return keys


def sat_period(pixel_length):
    """The pixel length in table cells when it is an integer (the phase-table
    period), else 1."""
    ratio = pixel_length / RES
    p = int(round(ratio))
    if 1 <= p <= 4096 and abs(ratio - p) < 1e-6:
        return p
    return 1


def layout_period(pixel_length):
    """SAT layout period: the pixel length in cells rounded to an integer.
    Pixel edges fall floor(ratio) or ceil(ratio) cells apart; with this period
    the more frequent step keeps a corner's residue, so the corners of
    neighbouring pixels mostly lie next to each other in memory (runs of ~2.6
    corners per cache line at 66.4 cells instead of one)."""
    ratio = pixel_length / RES
    p = int(round(ratio))
    return p if 2 <= p <= 4096 else 1


class PSFTables:
    """Per-depth-key SATs for one fluorophore/PSF model on one device."""

    def __init__(self, fluorophore_type, wavelength, radial_cutoff, depth_cutoff,
                 radial_width=None, pixel_length=None, device=None, na=1.4):
        self.device = device or _device.require_gpu()
        self.gaussian = fluorophore_type == 'Gaussian'
        if self.gaussian and radial_width is None:
            raise ValueError('fluorophore.radial_width must be given for Gaussian type fluorophore.')
        self.wavelength = wavelength
        self.radial_width = radial_width
        self.depth_cutoff = depth_cutoff
        self.nr = len(numpy.arange(0.0, radial_cutoff, RES, dtype=float))
        self.ntab = 2 * self.nr - 1
        self.period = layout_period(pixel_length) if pixel_length else 1
        self.sat_elems = int(_lib.lib().smx_psf_sat_elems(self.nr, self.period))
        # NA is 1.4 in the reference (_epifm.py:183); alpha, gamma as at :186-188
        k = 2.0 * numpy.pi / wavelength
        self.alpha = k * na
        self.gamma = k * numpy.power(na / 2, 2)
        self.scale = self.alpha * self.alpha / numpy.pi
        # key index: 0..n_keys-2 for depth keys (int(|d| / res) for |d| < cutoff + res,
        # with one key of margin), n_keys-1 for the cutoff key
        self.n_keys = int(round(depth_cutoff / RES)) + 3
        self.cutoff_index = self.n_keys - 1
        self.slot_of = {}
        self.capacity = 0
        self.sat = None
        # phase tables (integer pixel/cell ratio only): footprint pixels per axis
        self.m = int(math.ceil(RES * (self.ntab - 1) / pixel_length)) + 2 if pixel_length else 0
        self.use_phase = bool(pixel_length) and sat_period(pixel_length) > 1
        self.k_elems = (int(_lib.lib().smx_phase_table_elems(self.period, self.m))
                        if self.use_phase else 0)
        self.K = None
        self._K32 = None
        self._Kfold = None
        self.fold_elems = (int(_lib.lib().smx_phase_fold_elems(self.ntab, self.period))
                           if self.use_phase else 0)

    def key_index(self, key):
        return self.cutoff_index if key < 0 else int(key)

    def _depth_of(self, key):
        return self.depth_cutoff if key < 0 else key * RES

    def _grow(self, need):
        if need <= self.capacity:
            return
        cap = max(need, 2 * self.capacity, 1)
        new = torch.empty((cap, self.sat_elems), dtype=torch.float64, device=self.device)
        if self.sat is not None and self.capacity:
            new[: self.capacity].copy_(self.sat[: self.capacity])
        if self.use_phase:
            newk = torch.empty((cap, self.k_elems), dtype=torch.float64, device=self.device)
            if self.K is not None and self.capacity:
                newk[: self.capacity].copy_(self.K[: self.capacity])
            self.K = newk
        self.sat, self.capacity = new, cap

    def ensure(self, keys):
        """Build the tables for every key in ``keys`` not built yet."""
        if self.gaussian:
            keys = [0]   # the Gaussian PSF does not depend on depth (_epifm.py:133)
        missing = sorted({int(k) for k in numpy.unique(keys)} - set(self.slot_of))
        if not missing:
            return
        base = len(self.slot_of)
        self._grow(base + len(missing))
        lib, st = _lib.lib(), _device.stream()
        nz = len(missing)
        radial = torch.empty((nz, self.nr), dtype=torch.float64, device=self.device)
        if self.gaussian:
            _lib.call('smx_psf_radial_gaussian', self.nr, RES, self.radial_width,
                      radial.data_ptr(), st)
        else:
            depths = _device.upload(numpy.array([self._depth_of(k) for k in missing]),
                                    device=self.device)
            j0w = _device.ARENA.typed('psf_j0', self.nr * 100, torch.float64, self.device)
            _lib.call('smx_psf_radial_born_wolf', depths.data_ptr(), nz, self.nr, RES, self.alpha,
                      self.gamma, self.scale, j0w.data_ptr(), radial.data_ptr(), st)
        # build in batches so the row-sum scratch stays bounded (~64 MB per table)
        batch = 8
        for s in range(0, nz, batch):
            m = min(batch, nz - s)
            ws = _device.ARENA.get('psf_satws', lib.smx_psf_sat_workspace_bytes(self.nr, m),
                                   self.device)
            _lib.call('smx_psf_sat_build', radial[s:s + m].data_ptr(), m, self.nr, RES,
                      self.period, ws.data_ptr(), self.sat[base + s].data_ptr(), st)
        if self.use_phase:
            _lib.call('smx_phase_table_build', self.sat[base].data_ptr(), self.sat_elems, nz,
                      self.ntab, self.period, self.m, RES * RES, self.K[base].data_ptr(), st)
        for i, k in enumerate(missing):
            self.slot_of[k] = base + i
        self.radial_last = radial

    def k32(self):
        """float32 copy of the phase tables (rebuilt after new keys)."""
        if not self.use_phase:
            return None
        n = len(self.slot_of) * self.k_elems
        if self._K32 is None or self._K32.numel() != n:
            self._K32 = torch.empty(n, dtype=torch.float32, device=self.device)
            _lib.call('smx_phase_table_to_float', self.K.data_ptr(), n, self._K32.data_ptr(),
                      _device.stream())
        return self._K32

    def kfold(self):
        """Folded float32 phase tables: one block per unordered pair of
        canonical phases (about 1/8 of the full table; rebuilt after new keys)."""
        if not self.use_phase:
            return None
        nslots = len(self.slot_of)
        n = nslots * self.fold_elems
        if self._Kfold is None or self._Kfold.numel() != n:
            self._Kfold = torch.empty(n, dtype=torch.float32, device=self.device)
            _lib.call('smx_phase_table_fold', self.K.data_ptr(), nslots, self.ntab, self.period,
                      self.m, self._Kfold.data_ptr(), _device.stream())
        return self._Kfold

    def slot_map(self):
        """Device int32 table: key index -> slot (-1 if not built)."""
        m = numpy.full(self.n_keys, -1, dtype=numpy.int32)
        for k, s in self.slot_of.items():
            m[self.key_index(k)] = s
        if self.gaussian:
            m[:] = self.slot_of.get(0, -1)
        return _device.upload(m, dtype=torch.int32, device=self.device)

    def prebuild_all(self):
        """Build every depth key (0 .. cutoff) plus the cutoff key."""
        keys = list(range(self.n_keys - 2)) + [-1]
        self.ensure(keys)

    def prebuild_range(self, max_abs_depth):
        """Build the keys of every |depth| <= max_abs_depth (from the focal
        plane) plus the cutoff key."""
        top = min(self.n_keys - 2, int(max_abs_depth / RES) + 2)
        self.ensure(list(range(top)) + [-1])


class LUTTables:
    """Radial x depth lookup-table PSF on one device ("LUT mode").

    A float64 radial profile on `nb` bins over [0, r_max] for each of `nz`
    depth slices evenly spaced over [0, depth cutoff] (Born-Wolf,
    _epifm.py:180-213; the Gaussian profile, which does not depend on depth,
    takes one slice), and per slice the float32 pixel integrals of the
    reference's cartesian kernel (the profile clamped at r_max on the square
    |x|, |y| <= r_max, _epifm.py:98-126) at (q + 1)^2 sub-pixel phases
    (smx_lut_build, `K`), and their cell form for the splat (`cells`,
    smx_lut_cells: per pixel the four phases around each of q^2 cells).  Emitters take the nearest slice of their 1 nm depth key
    and the bilinear combination of the four phases around their position.

    Compared with the reference's 1 nm depth keys and 1 nm cell sums this is
    a deliberate approximation: depth to the nearest of `nz` slices (half a
    slice spacing, 1.96 nm at 256), sub-pixel position to bilinear
    interpolation over 1/q of a pixel, and no rounding of pixel edges to the
    1 nm cell grid (which the reference does per emitter).
    """

    def __init__(self, fluorophore_type, wavelength, radial_cutoff, depth_cutoff,
                 radial_width=None, pixel_length=None, device=None, nz=256, nb=4096, q=32, ns=2,
                 na=1.4, build_bytes=1 << 29):
        self.device = device or _device.require_gpu()
        self.gaussian = fluorophore_type == 'Gaussian'
        if self.gaussian and radial_width is None:
            raise ValueError('fluorophore.radial_width must be given for Gaussian type fluorophore.')
        if not pixel_length:
            raise ValueError('LUT mode needs the pixel length')
        nr = len(numpy.arange(0.0, radial_cutoff, RES, dtype=float))
        self.r_max = (nr - 1) * RES                  # the reference's largest radial sample
        self.depth_cutoff = depth_cutoff
        self.nz = 1 if self.gaussian else int(nz)
        self.nb = int(nb)
        self.pixel_length = pixel_length
        c = int(math.floor(self.r_max / pixel_length)) + 1
        self.lut = _lib.LutT(int(q), 2 * c + 1, c, self.nz, int(ns), 0, float(pixel_length),
                             float(self.r_max))
        self.m = self.lut.m
        lib, st = _lib.lib(), _device.stream()
        self.slice_elems = int(lib.smx_lut_table_elems(self.lut))
        if self.slice_elems == 0:
            raise ValueError('invalid LUT description (q={}, pixel length {})'.format(q, pixel_length))
        self.slice_depths = numpy.linspace(0.0, depth_cutoff, self.nz) if self.nz > 1 else \
            numpy.zeros(1)
        dr = self.r_max / (self.nb - 1)
        self.radial = torch.empty((self.nz, self.nb), dtype=torch.float64, device=self.device)
        if self.gaussian:
            _lib.call('smx_psf_radial_gaussian', self.nb, dr, radial_width,
                      self.radial.data_ptr(), st)
        else:
            k = 2.0 * numpy.pi / wavelength
            alpha, gamma = k * na, k * numpy.power(na / 2, 2)
            depths = _device.upload(self.slice_depths, device=self.device)
            j0w = torch.empty(self.nb * 100, dtype=torch.float64, device=self.device)
            _lib.call('smx_psf_radial_born_wolf', depths.data_ptr(), self.nz, self.nb, dr, alpha,
                      gamma, alpha * alpha / numpy.pi, j0w.data_ptr(), self.radial.data_ptr(), st)
        per = int(lib.smx_lut_build_workspace_bytes(self.lut, 1))
        batch = max(1, min(self.nz, build_bytes // per))
        ws = torch.empty(per * batch, dtype=torch.uint8, device=self.device)
        self.K = torch.empty(self.nz * self.slice_elems, dtype=torch.float32, device=self.device)
        _lib.call('smx_lut_build', self.radial.data_ptr(), self.nz, self.nb, self.lut,
                  ws.data_ptr(), ws.numel(), self.K.data_ptr(), st)
        del ws
        # the splat's form: the four bilinear taps of a pixel in one 16-byte entry
        self.cell_elems = int(lib.smx_lut_cell_elems(self.lut))
        self.cells = torch.empty(self.nz * self.cell_elems, dtype=torch.float32,
                                 device=self.device)
        _lib.call('smx_lut_cells', self.K.data_ptr(), self.nz, self.lut, self.cells.data_ptr(), st)
        # key index -> slice: 0..n_keys-2 are 1 nm depth keys, n_keys-1 the cutoff key
        self.n_keys = int(round(depth_cutoff / RES)) + 3
        self.cutoff_index = self.n_keys - 1

    def slice_of_key(self, key_index):
        if self.nz == 1:
            return 0
        if key_index == self.cutoff_index:
            return self.nz - 1
        dz = self.depth_cutoff / (self.nz - 1)
        return int(min(self.nz - 1, round(key_index * RES / dz)))

    def slot_map(self):
        """Device int32 table: key index -> depth slice."""
        m = numpy.array([self.slice_of_key(k) for k in range(self.n_keys)], dtype=numpy.int32)
        return _device.upload(m, dtype=torch.int32, device=self.device)
