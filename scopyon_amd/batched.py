"""Many fields of view (FOVs) simulated together on one GPU.

Each FOV is an independent camera frame (its own size and emitter count) that
shares the optics/detector configuration.  One time step of the whole batch is
one fixed launch sequence -- Philox Brownian step with per-FOV walls, emitter
weights, tile-binned splat over every frame, Philox EMCCD + ADC over every
pixel -- so it can be captured once in a HIP graph and replayed
(``capture()``/``replay()``).  Philox counters live in device memory and advance
inside the graph, so every replay draws fresh numbers.

Layout in HBM (structure of arrays): emitter coordinates ``depth, u, v``
(float64, concatenated FOV after FOV, CSR ``emitter_offsets``), ``fov_id``
(int32 per emitter), per-emitter ``D``; frames concatenated at
``pix_off`` offsets: photons/expectation (float64) and ADC counts (float32 by
default).  The PSF summed-area tables are shared by all FOVs.

The splat works on square tiles of 16, 32 or 64 pixels, one wave per 16 x 16
or 32 x 32 quadrant; with ``tile_side='auto'`` (the default when the pixel is
an integer number of PSF cells, or in LUT mode) the constructor times the five
shapes on the batch it was given and keeps the fastest (``tile_timings``,
``autotune_tiles()``).  Frames do not depend on the tile shape.
"""
import os

import numpy
import torch

from . import _device, _lib
from ._epifm import EPIFMConfigs, FrameGeometry, Splatter, tiles_per_emitter
from ._psf import LUTTables, PSFTables

__all__ = ['BatchedSimulator']


class BatchedSimulator:

    def __init__(self, config, fov_shapes, n_emitters, *, D=0.1e-12, dt=None, seed=0,
                 boundary='reflect', box=None, depth_range=None, method='default',
                 adc_dtype=torch.float32, keep_expectation=False, table_precision='folded',
                 tile_side='auto', quadrant=16, fuse_detector=True, keep_photons=False,
                 psf='table',
                 lut_slices=256, lut_bins=4096, lut_phases=32):
        self.device = _device.require_gpu()
        conf = config[method] if hasattr(config, '__getitem__') else config
        self.configs = c = EPIFMConfigs(conf, rng=numpy.random.RandomState(seed))
        if c.filter_error is not None:
            raise TypeError(c.filter_error)
        if c.detector_type not in ('EMCCD', 'CCD', 'CMOS'):
            raise RuntimeError("Unknown detector type was given [{}]. ".format(c.detector_type)
                               + "Use either one of 'CMOS', 'CCD' or 'EMCCD'.")
        self.dt = float(dt if dt is not None else c.detector_exposure_time)
        self.seed = int(seed)
        self.nfov = len(fov_shapes)
        counts = numpy.broadcast_to(numpy.asarray(n_emitters, dtype=numpy.int64), (self.nfov,))
        self.emitter_offsets = numpy.concatenate([[0], numpy.cumsum(counts)])
        self.n = int(self.emitter_offsets[-1])
        pl = c.pixel_length
        p0 = c.detector_focal_point
        self.fov_shapes = [(int(nw), int(nh)) for nw, nh in fov_shapes]
        self.tile_side = 32
        self.quadrant = 16
        self._set_geometry(32)
        geoms = self.geoms
        if boundary not in ('none', 'periodic', 'reflect'):
            raise ValueError("boundary must be 'none', 'periodic' or 'reflect'")
        if psf not in ('table', 'lut'):
            raise ValueError("psf must be 'table' (1 nm depth keys) or 'lut'")
        self.lut = psf == 'lut'
        if self.lut:
            # radial x depth lookup table: every depth at once, any pixel size
            self.tables = LUTTables(c.fluorophore_type, c.psf_wavelength, c.radial_cutoff,
                                    c.depth_cutoff, c.psf_radial_width, pl, self.device,
                                    nz=lut_slices, nb=lut_bins, q=lut_phases)
        else:
            # PSF tables: depth 0 only for 2D batches, every key for 3D batches
            self.tables = PSFTables(c.fluorophore_type, c.psf_wavelength, c.radial_cutoff,
                                    c.depth_cutoff, c.psf_radial_width, pl, self.device)
            if depth_range is None:
                self.tables.ensure([0])
            elif boundary == 'none':
                # nothing keeps an emitter inside the depth range: every key up
                # to the cutoff, so no emitter can meet an unbuilt table
                self.tables.prebuild_all()
            else:
                # every 1 nm key an emitter inside the depth range (walls or
                # the periodic wrap keep it there) can reach, plus the cutoff key
                dmax = max(abs(depth_range[0]), abs(depth_range[1]))
                self.tables.prebuild_range(dmax)
        self.slot_map = self.tables.slot_map()
        self.splatter = Splatter(c, self.tables, self.device)
        self.det = c.detector_params()
        # fixed-pattern ADC offsets (_epifm.py:926-963): one map per FOV, drawn
        # FOV after FOV at that FOV's size from one RandomState(seed)
        self.gain_map = self.offset_map = None
        if c.ADConverter_fpn_type != 'none':
            fpn_rng = numpy.random.RandomState(self.seed)
            maps = [c._adc_gain(c.ADConverter_offset0, fpn_rng, shape) for shape in self.fov_shapes]
            self.offset_map = _device.upload(numpy.concatenate([o.ravel() for o, _ in maps]),
                                             device=self.device)
            self.gain_map = _device.upload(numpy.concatenate([g.ravel() for _, g in maps]),
                                           device=self.device)
        self._table = None
        if c.detector_type == 'CMOS':
            t = numpy.loadtxt(os.path.join(os.path.dirname(os.path.abspath(__file__)), 'catalog',
                                           'detector', 'RNDist_F40.csv'), delimiter=',')
            p = t[:, 1] / t[:, 1].sum()
            cdf = p.cumsum()
            cdf /= cdf[-1]
            self._table = (_device.upload(t[:, 0], device=self.device),
                           _device.upload(cdf, device=self.device), len(cdf))
        self.boundary = {'none': 0, 'periodic': 1, 'reflect': 2}[boundary]
        # emitters: uniform in each FOV's box (default: the frame extent)
        hostrng = numpy.random.Generator(numpy.random.Philox(self.seed))
        depth = numpy.zeros(self.n)
        u = numpy.empty(self.n)
        v = numpy.empty(self.n)
        boxes = numpy.zeros((self.nfov, 6))
        fov_id = numpy.empty(self.n, dtype=numpy.int32)
        for f, g in enumerate(geoms):
            a, b = self.emitter_offsets[f], self.emitter_offsets[f + 1]
            if box is None:
                hu, hv = g.fov.half_w, g.fov.half_h
                lo = [p0[0], p0[1] - hu, p0[2] - hv]
                hi = [p0[0], p0[1] + hu, p0[2] + hv]
            else:
                lo = [p0[0], p0[1] - box[0] / 2, p0[2] - box[1] / 2]
                hi = [p0[0], p0[1] + box[0] / 2, p0[2] + box[1] / 2]
            if depth_range is not None:
                lo[0], hi[0] = p0[0] + depth_range[0], p0[0] + depth_range[1]
                depth[a:b] = hostrng.uniform(lo[0], hi[0], b - a)
            else:
                depth[a:b] = p0[0]
            u[a:b] = hostrng.uniform(lo[1], hi[1], b - a)
            v[a:b] = hostrng.uniform(lo[2], hi[2], b - a)
            boxes[f] = lo + hi
            fov_id[a:b] = f
        self.ndim = 3 if depth_range is not None else 2
        self.depth = _device.upload(depth, device=self.device)
        self.u = _device.upload(u, device=self.device)
        self.v = _device.upload(v, device=self.device)
        self.fov_id = _device.upload(fov_id, dtype=torch.int32, device=self.device)
        # walls in (depth, u, v) order for the 3D step, (u, v) for the 2D one
        if self.ndim == 2:
            z = numpy.zeros(self.nfov)
            boxes = numpy.column_stack([boxes[:, 1], boxes[:, 2], z, boxes[:, 4], boxes[:, 5], z])
        self.fov_box = _device.upload(numpy.ascontiguousarray(boxes), device=self.device)
        if numpy.ndim(D) == 0:
            self.D = None
            self.D_scalar = float(D)
        else:
            self.D = _device.upload(numpy.asarray(D, dtype=float), device=self.device)
            self.D_scalar = 0.0
        self.p_state = None
        self._w_fresh = False
        # frame buffers and device counters
        self.photons = torch.zeros(self.n_pix, dtype=torch.float64, device=self.device)
        self.adc = torch.zeros(self.n_pix, dtype=adc_dtype, device=self.device)
        self.E = (torch.zeros(self.n_pix, dtype=torch.float64, device=self.device)
                  if keep_expectation else None)
        self.w = torch.empty(self.n, dtype=torch.float64, device=self.device)
        self.key = torch.empty(self.n, dtype=torch.int32, device=self.device)
        self.counter = torch.zeros(1, dtype=torch.int32, device=self.device)
        # tiles of side >= 16: at most tiles_per_emitter(..., 16) per emitter
        per = ((self.tables.m // 16 + 2) ** 2 if self.lut
               else tiles_per_emitter(pl, self.tables.ntab))
        self.capacity = max(1, self.n * per)
        self._set_workspace()
        self.use_phase = (not self.lut) and self.tables.use_phase
        # 'float64': the phase tables themselves (same values as the SAT path);
        # 'float32': their float32 copy (half the table traffic); 'folded': the
        # symmetric float32 fold (1/8 of the table, L2-resident).  With float32
        # each pixel contribution is a float32 product of the table value and
        # the emitter weight (relative error <= ~3 x 2^-24), summed in float64.
        if table_precision not in ('float64', 'float32', 'folded'):
            raise ValueError('table_precision must be float64, float32 or folded')
        self.k_format = 0
        self.Ktab = None
        if self.lut:
            self.Ktab = self.tables.cells
        if self.use_phase:
            self.k_format = {'float64': 0, 'float32': 1, 'folded': 2}[table_precision]
            if self.k_format == 0:
                self.Ktab = self.tables.K
            elif self.k_format == 1:
                self.Ktab = self.tables.k32()
            else:
                self.Ktab = self.tables.kfold()
        # EMCCD frames on phase tables: the detector runs inside the tile pass
        # (smx_splat_detect_ts), the photon image stays on chip unless
        # keep_photons asks for it
        self.fused = (bool(fuse_detector) and (self.use_phase or self.lut) and
                      c.detector_type == 'EMCCD')
        self.keep_photons = bool(keep_photons) or not self.fused
        self.graph = None
        self.graph_steps = 1
        self.tile_timings = {}
        if tile_side == 'auto':
            if self.use_phase or self.lut:
                self.autotune_tiles()
        else:
            shape = (int(tile_side), int(quadrant))
            if shape != (32, 16):
                if shape not in self.TILE_SHAPES or not (self.use_phase or self.lut):
                    raise ValueError('tile_side must be 16, 32 or 64 with quadrant 16, or 32 or '
                                     '64 with quadrant 32 (32 / 16 without phase tables)')
                self.quadrant = shape[1]
                self._set_geometry(shape[0])
                self._set_workspace()

    # ---- tiles ----
    TILE_SIDES = (16, 32, 64)
    # (tile side, quadrant side): q^2 waves per tile, one per quadrant
    TILE_SHAPES = ((16, 16), (32, 16), (64, 16), (32, 32), (64, 32))

    @staticmethod
    def shape_name(side, quadrant):
        return str(side) if quadrant == 16 else '{}/q{}'.format(side, quadrant)

    def _tile_arg(self):
        """The C ABI's tile argument: the side, with SMX_TILE_Q32 for 32 x 32
        quadrants and SMX_TILE_WARM once this workspace has been binned with
        the current layout (its counters are back at zero, its tile map is
        current: the binning skips their initialisation)."""
        return self.tile_side | (0x100 if self.quadrant == 32 else 0) | (0x200 if self._warm else 0)

    def _set_geometry(self, ts):
        """FOV descriptors counted in tiles of side ts."""
        pl, p0 = self.configs.pixel_length, self.configs.detector_focal_point
        geoms, pix, tiles = [], 0, 0
        for (nw, nh) in self.fov_shapes:
            g = FrameGeometry(nw, nh, pl, p0[1], p0[2], pix_off=pix, tile_off=tiles, tile=ts)
            geoms.append(g)
            pix += nw * nh
            tiles += g.n_tiles
        self.geoms = geoms
        self.n_pix, self.n_tiles = pix, tiles
        self.pix_offsets = numpy.array([g.fov.pix_off for g in geoms] + [pix])
        raw = b''.join(bytes(g.fov) for g in geoms)
        self.fovs_dev = torch.from_numpy(numpy.frombuffer(raw, dtype=numpy.uint8).copy()).to(
            self.device)
        self.tile_side = ts
        self._warm = False

    def _set_workspace(self):
        self._warm = False
        self.ws = torch.empty(
            _lib.lib().smx_splat_workspace_bytes(self.n, self.n_tiles, self.capacity),
            dtype=torch.uint8, device=self.device)

    def autotune_tiles(self, reps=3):
        """Pick the splat tile shape -- side 16, 32 or 64 pixels with one wave
        per 16 x 16 quadrant, or side 32 or 64 with one wave per 32 x 32
        quadrant -- by timing what a step runs on this batch's current
        positions (binning + tile pass, with the detector epilogue when it is
        fused) with HIP events; a shape whose tile lists are too long to sort
        (status bit 1) is taken only if every shape has such lists.  Returns
        the (side, quadrant) kept; the timings (ms) stay in `tile_timings`,
        keyed by shape_name()."""
        if self.graph is not None:
            raise RuntimeError('autotune_tiles() before capture()')
        self.weights()
        run = self.splat_detect if self.fused else self.splat
        results = {}
        for shape in self.TILE_SHAPES:
            self.quadrant = shape[1]
            self._set_geometry(shape[0])
            self._set_workspace()
            self.splatter.status.zero_()
            run()                                  # warm (code objects, caches)
            torch.cuda.synchronize()
            unsorted = bool(int(self.splatter.status.item()) & 1)
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            for _ in range(reps):
                run()
            b.record()
            torch.cuda.synchronize()
            results[shape] = (unsorted, a.elapsed_time(b) / reps)
        self.splatter.status.zero_()
        self.tile_timings = {self.shape_name(*k): ms for k, (_, ms) in results.items()}
        best = min(results, key=lambda k: results[k])   # sorted lists first, then time
        self.quadrant = best[1]
        self._set_geometry(best[0])
        self._set_workspace()
        return best

    # ---- one step ----
    def move(self):
        st = _device.stream()
        xs = [self.depth, self.u, self.v] if self.ndim == 3 else [self.u, self.v, self.u]
        _lib.call('smx_brownian_philox', self.n, self.ndim, xs[0].data_ptr(), xs[1].data_ptr(),
                  xs[2].data_ptr(), _device.ptr(self.D), self.D_scalar, self.dt, self.boundary,
                  None, None, self.seed, 0, 0, self.fov_id.data_ptr(), self.nfov,
                  self.fov_box.data_ptr(), self.counter.data_ptr(), st)

    def form(self):
        """Noise-free expectation of every frame (photons per pixel before QE/background)."""
        self.weights()
        self.splat()

    def weights(self):
        """Per-emitter weight and depth-table slot (_epifm.py:1280-1315)."""
        _lib.call('smx_emitter_weights', self.n, self.depth.data_ptr(), None,
                  self.splatter.wparams, self.dt, None, None, self.slot_map.data_ptr(),
                  self.slot_map.numel(), self.w.data_ptr(), self.key.data_ptr(), None,
                  _device.stream())
        self._w_fresh = True

    def splat(self, entry='smx_splat_expected'):
        """Pixel integration of every FOV (entry: smx_splat_expected, or its two
        halves smx_splat_bin / smx_splat_accumulate), in tiles of side
        `tile_side`."""
        t = self.tables
        if self.lut:
            stages = {'smx_splat_expected': 3, 'smx_splat_bin': 1, 'smx_splat_accumulate': 2}
            _lib.call('smx_splat_lut_ts', self.n, self.u.data_ptr(), self.v.data_ptr(),
                      self.w.data_ptr(), self.key.data_ptr(), self.fov_id.data_ptr(),
                      self.fovs_dev.data_ptr(), self.nfov, self.n_tiles, t.cells.data_ptr(), t.lut,
                      self.capacity, 0, self.photons.data_ptr(), self.ws.data_ptr(),
                      self.ws.numel(), self.splatter.status.data_ptr(), _device.stream(),
                      stages[entry], self._tile_arg())
            self._warm |= bool(stages[entry] & 1)
            return
        _lib.call(entry + '_ts', self.n, self.u.data_ptr(), self.v.data_ptr(),
                  self.w.data_ptr(), self.key.data_ptr(), self.fov_id.data_ptr(),
                  self.fovs_dev.data_ptr(), self.nfov, self.n_tiles, t.sat.data_ptr(),
                  t.sat_elems, len(t.slot_of), t.ntab, t.period, 1e-9, 1e-18, self.capacity, 0,
                  _device.ptr(self.Ktab), self.k_format, t.m, self.photons.data_ptr(),
                  self.ws.data_ptr(), self.ws.numel(), self.splatter.status.data_ptr(),
                  _device.stream(), self._tile_arg())
        self._warm |= entry != 'smx_splat_accumulate'

    def detect(self):
        """EMCCD + ADC of every frame, then advance the frame counter."""
        self.emccd()
        self.advance()

    def advance(self):
        """Advance the device frame counter (the Philox frame of the next step)."""
        _lib.call('smx_counter_add', self.counter.data_ptr(), 1, _device.stream())

    def splat_detect(self, entry='smx_splat_detect_ts'):
        """Splat and EMCCD + ADC of every frame in one tile pass (the emitter
        weights must be current): same ADC as splat() then emccd(), without
        the photon image's round trip through memory (it is written only with
        keep_photons).  entry: smx_splat_detect_ts, or
        smx_splat_detect_accumulate_ts on an already binned workspace."""
        t = self.tables
        f32 = 1 if self.adc.dtype == torch.float32 else 0
        out = self.photons.data_ptr() if self.keep_photons else None
        if self.lut:
            stages = {'smx_splat_detect_ts': 3, 'smx_splat_detect_accumulate_ts': 2}
            _lib.call('smx_splat_lut_detect_ts', self.n, self.u.data_ptr(), self.v.data_ptr(),
                      self.w.data_ptr(), self.key.data_ptr(), self.fov_id.data_ptr(),
                      self.fovs_dev.data_ptr(), self.nfov, self.n_tiles, t.cells.data_ptr(), t.lut,
                      self.capacity, 0, out, self.ws.data_ptr(), self.ws.numel(),
                      self.splatter.status.data_ptr(), _device.stream(),
                      _device.ptr(self.gain_map), _device.ptr(self.offset_map), self.det,
                      self.seed, 0, 0, self.counter.data_ptr(), _device.ptr(self.E),
                      self.adc.data_ptr(), f32, stages[entry], self._tile_arg())
            self._warm |= bool(stages[entry] & 1)
            return
        _lib.call(entry, self.n, self.u.data_ptr(), self.v.data_ptr(), self.w.data_ptr(),
                  self.key.data_ptr(), self.fov_id.data_ptr(), self.fovs_dev.data_ptr(),
                  self.nfov, self.n_tiles, t.sat.data_ptr(), t.sat_elems, len(t.slot_of), t.ntab,
                  t.period, 1e-9, 1e-18, self.capacity, 0, _device.ptr(self.Ktab), self.k_format,
                  t.m, out, self.ws.data_ptr(), self.ws.numel(),
                  self.splatter.status.data_ptr(), _device.stream(), _device.ptr(self.gain_map),
                  _device.ptr(self.offset_map), self.det, self.seed, 0, 0,
                  self.counter.data_ptr(), _device.ptr(self.E), self.adc.data_ptr(), f32,
                  self._tile_arg())
        self._warm |= entry == 'smx_splat_detect_ts'

    def stages(self):
        """The launches of one step, in order, as callables (for timing)."""
        st = {'brownian_philox': self.move}
        if self.ndim == 3:
            st['emitter_weights'] = self.weights
        st['splat_bin'] = lambda: self.splat('smx_splat_bin')
        if self.fused:
            st['splat_detect'] = lambda: self.splat_detect('smx_splat_detect_accumulate_ts')
        else:
            st['splat_accumulate'] = lambda: self.splat('smx_splat_accumulate')
            st['emccd_philox'] = self.emccd
        st['counter_add'] = self.advance
        return st

    def emccd(self):
        """Detector + ADC of every pixel (EMCCD, or CCD / CMOS by configuration)."""
        f32 = 1 if self.adc.dtype == torch.float32 else 0
        if self.configs.detector_type == 'EMCCD':
            _lib.call('smx_emccd_philox', self.n_pix, self.photons.data_ptr(),
                      _device.ptr(self.gain_map), _device.ptr(self.offset_map), self.det,
                      self.seed, 0, 0, self.counter.data_ptr(), _device.ptr(self.E),
                      self.adc.data_ptr(), f32, _device.stream())
            return
        tv, tc, tn = self._table if self._table is not None else (None, None, 0)
        _lib.call('smx_photon_detector_philox', self.n_pix, self.photons.data_ptr(),
                  _device.ptr(self.gain_map), _device.ptr(self.offset_map), self.det,
                  _device.ptr(tv), _device.ptr(tc), tn, self.seed, 0, 0, self.counter.data_ptr(),
                  _device.ptr(self.E), self.adc.data_ptr(), f32, _device.stream())

    def set_positions(self, depth, u, v):
        """Overwrite the emitter coordinates (concatenated over FOVs)."""
        for dst, src in ((self.depth, depth), (self.u, u), (self.v, v)):
            dst.copy_(torch.as_tensor(numpy.ascontiguousarray(src, dtype=numpy.float64)))
        if self.ndim == 2:
            self.weights()          # depth is not moved in 2D: weights stay current

    def _step_weights(self):
        # 2D batches keep every emitter at its depth and the exposure fixed, so
        # the weights and table slots of the first step hold for every step
        if self.ndim == 3 or not self._w_fresh:
            self.weights()

    def step(self):
        """Advance every FOV by one time step and form + detect its frame."""
        self.move()
        self._step_weights()
        if self.fused:
            self.splat_detect()
            self.advance()
        else:
            self.splat()
            self.detect()

    # ---- HIP graph ----
    def capture(self, steps=1):
        """Capture `steps` consecutive steps in one HIP graph (static buffers,
        device counters); replay() then runs them `steps` at a time."""
        steps = int(steps)
        if steps < 1:
            raise ValueError('steps must be >= 1')
        self.step()          # warm: workspace, tables, code objects
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            for _ in range(steps):
                self.step()
        self.graph = g
        self.graph_steps = steps
        return g

    def replay(self, steps=1):
        """Advance `steps` time steps: graph replays when captured (steps must
        be a multiple of the captured step count), else eager steps."""
        if self.graph is None:
            for _ in range(steps):
                self.step()
            return
        if steps % self.graph_steps:
            raise ValueError('replay({}) with a graph of {} steps'.format(steps, self.graph_steps))
        for _ in range(steps // self.graph_steps):
            self.graph.replay()

    # ---- outputs ----
    def frame(self, f, which='adc'):
        buf = {'adc': self.adc, 'photons': self.photons, 'expectation': self.E}[which]
        a, b = self.pix_offsets[f], self.pix_offsets[f + 1]
        g = self.geoms[f].fov
        return buf[a:b].view(g.nw, g.nh)

    def status(self):
        """Status bits of the splats since construction: raises on a tile
        list overflow (bit 2) or an emitter that met an unbuilt PSF table
        (bit 8, its light would be missing); returns the word otherwise (bit
        1: a tile list of more than 8192 entries, summed out of emitter
        order)."""
        s = int(self.splatter.status.item())
        if s & 2:
            raise RuntimeError('splat tile list capacity exceeded')
        if s & 8:
            raise RuntimeError('a PSF table needed by an emitter was not built')
        return s
