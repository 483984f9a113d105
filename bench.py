"""Benchmark: batched TIRF/EMCCD frame simulation on one MI355X.

Workload (the metric's configuration): F fields of view, each a 512 x 512
EMCCD frame (16 um pixels, 100x magnification -> 160 nm) with 10,000 emitters
diffusing in 2D (D = 0.1 um^2/s, dt = 33 ms, reflecting walls at the frame
edge), TRITC Born-Wolf PSF integrated per pixel (phase tables built from the
float64 summed-area table), EM gain 300, readout sigma 100 e-, 16-bit ADC.  One
step = Brownian step + emitter weights + splat + EMCCD/ADC for every frame of
the batch, replayed from one captured HIP graph.  Inputs are synthetic and
resident in HBM.

Also reported: the plain single-frame path (same kernels, one frame per step,
eager launches), the drop-in form_image() path (RandomState stream,
reference-exact EMCCD), per-kernel times from HIP events, the roofline of the
dominant kernel (traffic from the committed rocprofv3 --pmc summary), the
differentiable forward+backward rate, and the CPU baseline (the float64 oracle
on this host, one core).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--fovs F]
    python bench.py --calibrate-copy      (copy kernel for counter scaling)
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("frames/sec (512², 10k emitters) batched vs plain at ≤1e-5 err; "
          "fwd+bwd/sec")
HBM_PEAK_GBS = 8000.0
# stage -> the kernel(s) it launches
KERNELS = {
    'brownian_philox': 'brownian_philox_kernel',
    'emitter_weights': 'weights_kernel',
    'splat_bin': 'emitter_rec_kernel + tiles_init / compact / scan + bin_fill_rec_kernel',
    'splat_accumulate': 'splat_phase_kernel',
    'splat_detect': 'splat_phase_kernel (detector epilogue)',
    'emccd_philox': 'emccd_queue_kernel',
    'counter_add': 'counter_add_kernel',
}


def _progress(msg):
    """One progress line on stderr per finished stage of the benchmark."""
    print('[bench {:7.1f} s] {}'.format(time.perf_counter() - _T0, msg), file=sys.stderr,
          flush=True)


_T0 = time.perf_counter()


def _config(size, exposure):
    import scopyon_amd
    cfg = scopyon_amd.DefaultConfiguration()
    cfg.update({'default': {'magnification': 100.0,
                            'detector': {'image_size': [size, size], 'exposure_time': exposure},
                            'effects': {'photo_bleaching': {'switch': False}}}})
    return cfg


def graph_times(stages, reps):
    """Average launch duration (ms) of each stage: `reps` back-to-back
    launches of it captured in one HIP graph, replayed once between two HIP
    events on the stream the graph runs on (torch's current stream, where the
    kernels are launched).  No host launch gaps are counted."""
    import torch
    out = {}
    for name, fn in stages.items():
        fn()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            for _ in range(reps):
                fn()
        g.replay()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        g.replay()
        b.record()
        torch.cuda.synchronize()
        out[name] = a.elapsed_time(b) / reps
        del g
    return out


def _kernel_times(sim, reps):
    """Per-stage launch durations (ms) of one batched step (graph_times)."""
    return graph_times(sim.stages(), reps)


def _stage_bytes(N, HW):
    """Algorithmic HBM bytes per launch of each stage (N emitters, HW pixels):
    (bytes by the measurement convention of the workload description, bytes
    this implementation moves at the least)."""
    return {
        # 2D float64 coordinates read + written (32 B) and the FOV id (4 B)
        'brownian_philox': (N * 36, N * 36),
        # depth read (8 B); weight (8 B) and table slot (4 B) written
        'emitter_weights': (N * 20, N * 20),
        # u, v, w read (24 B); one 40 B table record written
        'splat_bin': (N * 64, N * 64),
        # convention: 16 B of emitter data read, 4 B per pixel written; here:
        # position and weight (24 B) read, the float64 photon image (8 B/px) written
        'splat_accumulate': (N * 16 + HW * 4, N * 24 + HW * 8),
        # float64 photons read (8 B/px), float32 ADC written (4 B/px)
        'emccd_philox': (HW * 8, HW * 12),
        # splat + EMCCD + ADC in one pass: emitters in, float32 ADC out
        'splat_detect': (N * 16 + HW * 4, N * 24 + HW * 4),
    }


def _traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE, each scaled by the copy-kernel calibration), or None."""
    path = os.path.join(ROOT, 'profiles', 'pmc_traffic.json')
    try:
        with open(path) as f:
            data = json.load(f)
        return data['kernels'][kernel]['hbm_bytes_per_launch']
    except (OSError, KeyError, ValueError):
        return None


def calibrate_copy(n_bytes=1 << 28, reps=5):
    """A plain device copy of known size, for scaling the FETCH_SIZE/WRITE_SIZE
    counters: each launch reads and writes n_bytes."""
    import torch
    a = torch.ones(n_bytes // 8, dtype=torch.float64, device='cuda')
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({'copy_bytes_read': n_bytes, 'copy_bytes_written': n_bytes, 'reps': reps}))


def fwd_bwd(reps=20, warmup=3):
    """Differentiable forward + backward (separable Gaussian PSF, sigma 1.2 px):
    64 frames of 256 x 256, 5,000 emitters each, z ~ U[-0.5, 0.5] um, Poisson NLL
    against a seeded synthetic EMCCD frame (Philox EMCCD on the true image,
    converted back to photo-electrons).  One iteration = forward (binning +
    tiles on the matrix cores) with the NLL and dL/dmu in the tile pass's
    epilogue + backward for x, y, z, I and sigma, replayed from a HIP graph
    (the separate forward, NLL, backward sequence is timed beside it).
    Returns rates, per-stage times (graph_times) and the HBM roofline of the
    iteration (algorithmic bytes 52 N + 12 HW per frame)."""
    import ctypes
    import numpy
    import torch
    from scopyon_amd import _device, _lib
    from scopyon_amd.gauss import GaussianModel
    B, S, N = 64, 256, 5000
    rng = numpy.random.Generator(numpy.random.Philox(7))
    n = B * N
    dev = lambda a, t=torch.float32: torch.as_tensor(a, dtype=t, device='cuda')
    tx, ty = rng.uniform(0, S, n), rng.uniform(0, S, n)
    z = dev(rng.uniform(-0.5, 0.5, n))
    I = dev(rng.uniform(500.0, 1500.0, n))
    fid = dev(numpy.repeat(numpy.arange(B), N), torch.int32)
    model = GaussianModel([(S, S)] * B, sigma=1.2, inv_depth=1.0 / 0.377, background=10.0,
                          radius=6, method='mfma-f32')
    mu_true = model.forward(dev(tx), dev(ty), z, I, fid)
    det = _lib.DetectorParamsT()
    det.qe, det.background, det.background_switch, det.emgain = 0.92, 0.0, 0, 300
    det.readout_noise, det.fullwell, det.adc_max = 100.0, 800000.0, 65535.0
    det.offset, det.gain = 2000.0, 800000.0 / 63536.0
    adc = torch.empty(model.n_pix, dtype=torch.float32, device='cuda')
    photons = mu_true.double()
    _lib.call('smx_emccd_philox', model.n_pix, photons.data_ptr(), None, None, ctypes.byref(det),
              7, 0, 0, None, None, adc.data_ptr(), 1, _device.stream())
    k = ((adc - det.offset) * det.gain / det.emgain).clamp(min=0.0) / det.qe
    x, y = dev(tx + rng.normal(0, 0.3, n)), dev(ty + rng.normal(0, 0.3, n))
    mu = torch.empty(model.n_pix, dtype=torch.float32, device='cuda')
    grad = torch.empty_like(mu)
    outs = [torch.empty(n, dtype=torch.float32, device='cuda') for _ in range(4)]

    def fwd():
        model.forward(x, y, z, I, fid, out=mu)

    def nll():
        model.nll(mu, k, grad)

    def fwd_nll():
        model.forward_nll(x, y, z, I, k, fid, grad_mu=grad)

    def bwd():
        model.backward(x, y, z, I, grad, fid, outputs=outs)

    def it():
        fwd_nll()
        bwd()

    def it_separate():
        fwd()
        nll()
        bwd()
    for _ in range(warmup):
        it()
    torch.cuda.synchronize()
    it_separate()
    times = graph_times({'forward_nll': fwd_nll, 'backward': bwd, 'forward': fwd,
                         'nll': nll}, 10)

    def rate(fn):
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            fn()
        g.replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            g.replay()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps
    sec_sep = rate(it_separate)
    sec = rate(it)
    model.check_status()
    alg = B * (52 * N + 12 * S * S)
    return {'workload': '64 frames 256x256, 5000 emitters each, Gaussian PSF sigma 1.2 px, '
                        'forward on the matrix cores (float32 operands), Poisson NLL, '
                        'backward for x, y, z, I, sigma; hipGraph replay',
            'fwd_bwd_per_sec': 1.0 / sec, 'frames_per_sec': B / sec, 'ms_per_iteration': sec * 1e3,
            'ms_per_iteration_separate_nll': sec_sep * 1e3,
            'stage_ms': times, 'loss': float(model.loss.sum()),
            'roofline': {'bound': 'hbm', 'achieved': alg / sec / 1e9, 'peak': HBM_PEAK_GBS,
                         'unit': 'GB/s', 'frac': alg / sec / 1e9 / HBM_PEAK_GBS,
                         'algorithmic_bytes': alg}}


def precision_leg(reps=10):
    """Separable Gaussian PSF (sigma 1.2 px) over 1,000,000 emitters uniform on
    one 2048 x 2048 frame, intensities exponential (mean 500 photons): the
    contraction Gx^T diag(I) Gy on the matrix cores with bfloat16, float16 and
    float32 operands (float32 accumulation) and the float32 LDS splat.  Each
    form's time per frame (binning included) and its error against the float64
    oracle on a 256 x 256 crop: relative RMS ||mu - ref|| / ||ref|| and the
    largest relative error over pixels above 1% of the crop maximum."""
    import numpy
    import torch
    from oracle import gauss as OG
    from scopyon_amd.gauss import GaussianModel
    S, N, R, C = 2048, 1_000_000, 6, 256
    rng = numpy.random.Generator(numpy.random.Philox(11))
    x, y = rng.uniform(0, S, N), rng.uniform(0, S, N)
    I = rng.exponential(500.0, N)
    dev = lambda a: torch.as_tensor(a, dtype=torch.float32, device='cuda')
    xd, yd, Id, zd = dev(x), dev(y), dev(I), dev(numpy.zeros(N))
    x32, y32, I32 = (a.astype(numpy.float32).astype(float) for a in (x, y, I))
    # float64 oracle on the crop [c0, c0 + C)^2 (integer shifts keep G exact)
    c0 = S // 2 - C // 2
    sel = ((x32 >= c0 - R - 1) & (x32 < c0 + C + R + 1) & (y32 >= c0 - R - 1) &
           (y32 < c0 + C + R + 1))
    kw = dict(sigma=1.2, inv_depth=0.0, background=0.0, radius=R)
    ref = OG.render([(C, C)], x32[sel] - c0, y32[sel] - c0, numpy.zeros(sel.sum()), I32[sel],
                    numpy.zeros(sel.sum(), int), **kw)[0]
    big = ref > 0.01 * ref.max()
    # executed MFMA work: every tile list padded to 16 emitters
    lo_r = numpy.floor(x32).astype(int) - R
    lo_c = numpy.floor(y32).astype(int) - R
    r0, r1 = numpy.clip(lo_r, 0, S - 1) // 32, numpy.clip(lo_r + 2 * R, 0, S - 1) // 32
    q0, q1 = numpy.clip(lo_c, 0, S - 1) // 32, numpy.clip(lo_c + 2 * R, 0, S - 1) // 32
    counts = numpy.zeros((S // 32, S // 32), dtype=numpy.int64)
    for dr in (0, 1):
        for dc in (0, 1):
            ok = (r0 + dr <= r1) & (q0 + dc <= q1)
            numpy.add.at(counts, (r0[ok] + dr, q0[ok] + dc), 1)
    padded = int(numpy.sum((counts + 15) // 16 * 16))
    out = {'workload': '1,000,000 emitters on 2048x2048, Gaussian sigma 1.2 px, window 13 px, '
                       'I ~ Exp(500)',
           'dense_equivalent_flop': 2.0 * N * S * S, 'forms': {}}
    for method in ('mfma-bf16', 'mfma-f16', 'mfma-f32', 'lds'):
        m = GaussianModel([(S, S)], method=method, **kw)
        zt = zd if method == 'lds' else None
        mu = m.forward(xd, yd, zt, Id)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            m.forward(xd, yd, zt, Id, out=mu)
        b.record()
        torch.cuda.synchronize()
        ms = a.elapsed_time(b) / reps
        m.check_status()
        got = mu.view(S, S)[c0:c0 + C, c0:c0 + C].cpu().numpy().astype(float)
        rms = float(numpy.linalg.norm(got - ref) / numpy.linalg.norm(ref))
        mx = float((numpy.abs(got - ref)[big] / ref[big]).max())
        entry = {'ms_per_frame': ms, 'frames_per_sec': 1e3 / ms, 'rel_rms_err': rms,
                 'max_rel_err': mx}
        if method != 'lds':
            peak = {'mfma-bf16': 2.5e15, 'mfma-f16': 2.5e15, 'mfma-f32': 157.3e12}[method]
            flop = 2.0 * 32 * 32 * padded
            entry.update({'executed_mfma_flop': flop, 'peak_flops': peak,
                          'mfma_utilisation': flop / (ms * 1e-3) / peak,
                          'dense_equivalent_tflops': 2.0 * N * S * S / (ms * 1e-3) / 1e12})
        out['forms'][method] = entry
        del m
    return out


def _replay_rate(sim, steps, warmup, graph_steps=1):
    import torch
    sim.capture(graph_steps)
    for _ in range(warmup):
        sim.replay(graph_steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.replay(steps)
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / steps
    assert sim.status() & ~1 == 0
    return sec


def scale3d_leg(steps=20, warmup=2):
    """EPIFM 3D scale-up: 100,000 emitters uniform in a 50 x 50 x 2 um box,
    D lognormal (median 0.5 um^2/s, sigma_log 1.0), dt 10 ms, EPI
    illumination, 1024 x 1024 EMCCD frames at 241x (66.4 nm pixels).  Two PSF
    forms: the radial x depth lookup table (256 depth slices x 4096 radial
    bins, float64-built, pixel integrals at 33 x 33 sub-pixel phases,
    bilinear; detector fused into the tile pass) and the 1 nm parity tables
    (a float64 summed-area table for every 1 nm depth key of the +-1 um
    range, 1002 tables, 32 GB of HBM, general summed-area splat)."""
    import numpy
    import torch
    import scopyon_amd
    from scopyon_amd.batched import BatchedSimulator
    cfg = scopyon_amd.DefaultConfiguration()
    cfg.update({'default': {'detector': {'image_size': [1024, 1024], 'exposure_time': 0.01},
                            'light_source': {'angle': 0.0},
                            'effects': {'photo_bleaching': {'switch': False}}}})
    n = 100_000
    rng = numpy.random.Generator(numpy.random.Philox(13))
    D = numpy.exp(rng.normal(numpy.log(0.5e-12), 1.0, n))
    out = {'workload': '1 FOV 1024x1024 at 241x, 100,000 emitters in 50x50x2 um, lognormal D, '
                       'EPI, dt 10 ms'}
    for mode in ('lut', 'table'):
        t0 = time.perf_counter()
        sim = BatchedSimulator(cfg, [(1024, 1024)], n, D=D, dt=0.01, seed=21,
                               box=(50e-6, 50e-6), depth_range=(-1e-6, 1e-6), psf=mode)
        torch.cuda.synchronize()
        build = time.perf_counter() - t0
        sec = _replay_rate(sim, steps, warmup)
        entry = {'frames_per_sec': 1.0 / sec, 'ms_per_frame': sec * 1e3, 'table_build_s': build,
                 'tile_side': sim.tile_side, 'quadrant': sim.quadrant,
                 'kernel_ms': _kernel_times(sim, 5)}
        if mode == 'lut':
            t = sim.tables
            entry.update({'psf': 'radial x depth LUT: {} slices x {} radial bins (float64), '
                                 '{}^2 phases x {}^2 px float32 pixel integrals, bilinear '
                                 '(splat reads the four taps of a pixel as one 16-byte '
                                 'cell entry)'.format(t.nz, t.nb, t.lut.q + 1, t.m),
                          'table_bytes': t.K.numel() * 4,
                          'cell_table_bytes': t.cells.numel() * 4})
            out['lut'] = entry
        else:
            entry.update({'psf': '1 nm depth keys, float64 summed-area tables (parity)',
                          'depth_tables': len(sim.tables.slot_of)})
            out['parity_tables'] = entry
        del sim
        torch.cuda.empty_cache()
    out['frames_per_sec'] = out['lut']['frames_per_sec']
    return out


def batched_uneven_leg(steps=20, warmup=3):
    """Batched fields of view: 256 independent simulations, frame sizes drawn
    from {128^2, 256^2, 512^2}, emitter counts lognormal (median 5000,
    sigma_log 1.0) clipped to [100, 50000], Philox per FOV, one HIP graph per
    step."""
    import numpy
    from scopyon_amd.batched import BatchedSimulator
    rng = numpy.random.Generator(numpy.random.Philox(17))
    F = 256
    sizes = rng.choice([128, 256, 512], F)
    counts = numpy.clip(numpy.exp(rng.normal(numpy.log(5000.0), 1.0, F)), 100, 50000).astype(int)
    sim = BatchedSimulator(_config(512, 0.033), [(int(s), int(s)) for s in sizes], counts,
                           D=0.1e-12, dt=0.033, seed=23)
    sec = _replay_rate(sim, steps, warmup)
    # the same steps captured ten to a graph (one graph launch per 10 steps)
    sec10 = _replay_rate(sim, steps, warmup, graph_steps=10)
    return {'workload': '256 FOVs, sizes from {128,256,512}^2, lognormal emitter counts '
                        '(median 5000) clipped to [100, 50000], hipGraph replay',
            'fovs': F, 'emitters': int(counts.sum()), 'pixels': int(sim.n_pix),
            'frames_per_sec': F / sec, 'ms_per_step': sec * 1e3,
            'ms_per_step_10_steps_per_graph': sec10 * 1e3,
            'emitter_frames_per_sec': float(counts.sum()) / sec,
            'status_unsorted_lists': bool(sim.status() & 1),
            'tile_side': sim.tile_side, 'quadrant': sim.quadrant,
            'tile_autotune_ms': {str(k): v for k, v in sim.tile_timings.items()}}


def _cpu_model():
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return 'unknown'


def _oracle_frame_setup(size, n_emit, seed):
    import numpy
    from oracle import epifm as O, params as P
    ocfg = P.default_config()
    ocfg['magnification'] = 100.0
    ocfg['detector']['image_size'] = [size, size]
    ocfg['detector']['exposure_time'] = 0.033
    ocfg['effects']['photo_bleaching']['switch'] = False
    p = P.epifm_params(ocfg, numpy.random.RandomState(0))
    rng = numpy.random.RandomState(seed)
    half = 0.5 * size * (16e-6 / 100.0)
    xy = rng.uniform(-half, half, size=(n_emit, 2))
    parts = numpy.column_stack([numpy.zeros(n_emit), xy, numpy.arange(n_emit), numpy.ones(n_emit)])
    psf = O.PSFCache(p)
    psf.get(0.0)    # the table build is a one-time cost, not per frame
    return p, psf, parts, rng


def _pool_splat(args):
    from oracle import epifm as O
    p, table, parts, size = args
    psf = O.PSFCache(p)
    psf.tables[0] = table
    import numpy
    exp, _, _ = O.molecule_plane(p, psf, parts, (size, size), numpy.zeros(3), 0.033, {}, None,
                                 None)
    return exp


def _pool_emccd(args):
    import numpy
    from oracle import epifm as O
    E, emgain, seed = args
    return O.emccd_signal(E, emgain, numpy.random.RandomState(seed), memo=False)


def cpu_baseline_child(size, n_emit, seed=1, procs=None):
    """CPU timings of one headline frame (no GPU is touched in this process):
    1. the line-faithful float64 oracle on one core: all n_emit emitters
       splatted (_epifm.py:1227-1333) and the per-pixel EMCCD loop over every
       pixel (_epifm.py:402-416, pmf rebuilt per pixel as the reference does);
    2. the same oracle fanned out over `procs` processes (emitter chunks, then
       pixel chunks; the reference's processes > 1 Pool, _epifm.py:1239-1260,
       404-407; not stream-equal, as the reference's Pool uses the global RNG);
    3. a vectorised-numpy frame (oracle/vectorised.py: phase-table splat,
       numpy Generator EMCCD), the best-effort single-core CPU form."""
    import multiprocessing
    import numpy
    from oracle import epifm as O, vectorised as V
    p, psf, parts, rng = _oracle_frame_setup(size, n_emit, seed)
    t0 = time.perf_counter()
    exp, _, _ = O.molecule_plane(p, psf, parts, (size, size), numpy.zeros(3), 0.033, {}, None,
                                 None)
    t_splat = time.perf_counter() - t0
    t0 = time.perf_counter()
    E = p.QE * (exp + p.background_mean)
    noise = O.emccd_noise(E.shape, p.readout_noise, rng)
    sig = O.emccd_signal(E, p.emgain, rng, memo=False)
    O.adc_counts(sig + noise, p.adc_fullwell, p.adc_gain, p.adc_offset, p.adc_bit)
    t_det = time.perf_counter() - t0
    out = {'cpu_model': _cpu_model(), 'cpu_count': os.cpu_count(),
           'single_core': {'seconds_per_frame': t_splat + t_det,
                           'seconds': {'splat': t_splat, 'detector': t_det}}}
    procs = procs or min(16, os.cpu_count() or 1)
    table = psf.get(0.0)
    chunks = numpy.array_split(parts, procs)
    ctx = multiprocessing.get_context('fork')
    with ctx.Pool(procs) as pool:
        t0 = time.perf_counter()
        parts_img = pool.map(_pool_splat, [(p, table, c, size) for c in chunks])
        exp2 = numpy.sum(parts_img, axis=0)
        t_s2 = time.perf_counter() - t0
        t0 = time.perf_counter()
        E2 = p.QE * (exp2 + p.background_mean)
        noise2 = O.emccd_noise(E2.shape, p.readout_noise, rng)
        rows = numpy.array_split(E2, procs)
        sig2 = numpy.concatenate(pool.map(_pool_emccd, [(r, p.emgain, k) for k, r in
                                                          enumerate(rows)]))
        O.adc_counts(sig2 + noise2, p.adc_fullwell, p.adc_gain, p.adc_offset, p.adc_bit)
        t_d2 = time.perf_counter() - t0
    out['processes'] = {'procs': procs, 'seconds_per_frame': t_s2 + t_d2,
                        'seconds': {'splat': t_s2, 'detector': t_d2}}
    ps = V.PhaseSplat(psf.get(0.0), 16e-6 / 100.0)
    w = numpy.full(n_emit, p_weight(p))
    gen = numpy.random.Generator(numpy.random.Philox(seed))
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        img = ps.frame((size, size), parts[:, 1], parts[:, 2], w)
        V.emccd_adc(img, p, gen)
    t_v = (time.perf_counter() - t0) / reps
    out['vectorised'] = {'seconds_per_frame': t_v,
                         'max_rel_diff_vs_oracle': float(numpy.abs(img - exp).max() / exp.max())}
    return out


def p_weight(p):
    """Emitter weight at the focal plane (_epifm.py:1280-1315), from the
    oracle's parameters."""
    import numpy
    from oracle import params as P
    amp, _ = P.snells_law(p.light_source)
    n_emit = P.get_emit_photons(amp, 0.033, p.abs_coefficient, p.quantum_yield,
                                p.fluorophore_radius)
    return numpy.sum(p.fluoem_norm) * p.psf_normalization * (1.0 * n_emit / (4.0 * numpy.pi))


def cpu_baseline(size, n_emit, timeout=400):
    """Run cpu_baseline_child in a child process (so the CPU work never
    shares a process with the GPU runtime) and shape its timings into the
    benchmark's cpu_baseline record."""
    import subprocess
    cmd = [sys.executable, os.path.abspath(__file__), '--cpu-baseline-child', '--size', str(size),
           '--emitters', str(n_emit)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError('CPU baseline failed: ' + r.stderr[-2000:])
    d = json.loads(r.stdout.strip().splitlines()[-1])
    one = d['single_core']['seconds_per_frame']
    return {'value': 1.0 / one, 'unit': 'frames/s', 'cores': 1, 'kind': 'port',
            'sample': ('one full {0}x{0} frame, {1} emitters, z = 0: the float64 numpy/scipy '
                       'oracle, single core, splat of every emitter + the per-pixel EMCCD pmf '
                       'and inverse-CDF draw over all {2} pixels + ADC'.format(
                           size, n_emit, size * size)),
            'seconds': d['single_core']['seconds'],
            'cpu_model': d['cpu_model'], 'cpu_count': d['cpu_count'],
            'variants': {
                'processes': {'value': 1.0 / d['processes']['seconds_per_frame'],
                              'unit': 'frames/s', 'cores': d['processes']['procs'],
                              'seconds': d['processes']['seconds'],
                              'note': 'the oracle over a process pool (emitter chunks, then '
                                      'pixel chunks), like the reference with processes > 1'},
                'vectorised_numpy': {'value': 1.0 / d['vectorised']['seconds_per_frame'],
                                     'unit': 'frames/s', 'cores': 1,
                                     'max_rel_diff_vs_oracle':
                                         d['vectorised']['max_rel_diff_vs_oracle'],
                                     'note': 'phase-table splat + numpy Generator EMCCD '
                                             '(statistically equivalent detector)'}}}


def gpu_vs_oracle_error(size, sample_emitters=200, seed=1):
    """Max relative error of the batched path's noise-free expectation
    QE (photons + background) against the float64 oracle on a sample."""
    import numpy
    from oracle import epifm as O
    from scopyon_amd.batched import BatchedSimulator
    p, psf, parts, _ = _oracle_frame_setup(size, sample_emitters, seed)
    exp, _, _ = O.molecule_plane(p, psf, parts, (size, size), numpy.zeros(3), 0.033, {}, None,
                                 None)
    E = p.QE * (exp + p.background_mean)
    sim = BatchedSimulator(_config(size, 0.033), [(size, size)], sample_emitters, dt=0.033, seed=0)
    sim.set_positions(parts[:, 0], parts[:, 1], parts[:, 2])
    sim.form()
    got = sim.frame(0, 'photons').cpu().numpy()
    Eg = p.QE * (got + p.background_mean)
    return float(numpy.max(numpy.abs(Eg - E) / numpy.abs(E)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=50)
    ap.add_argument('--warmup', type=int, default=5)
    ap.add_argument('--fovs', type=int, default=64)
    ap.add_argument('--emitters', type=int, default=10000)
    ap.add_argument('--size', type=int, default=512)
    ap.add_argument('--plain-steps', type=int, default=20)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--legs', default='plain,api,uneven,scale3d,precision,fwd_bwd',
                    help='comma list of the extra legs to run besides the headline batch '
                         '(plain, api, uneven, scale3d, precision, fwd_bwd); "none" for only '
                         'the headline (for profiling runs)')
    ap.add_argument('--cpu-baseline-child', action='store_true', help=argparse.SUPPRESS)
    ap.add_argument('--calibrate-copy', action='store_true',
                    help='only run a device copy of known size (for rocprofv3 --pmc scaling)')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py measures one GPU only; --gpus {} is not supported (the FOV batch '
              'would shard across GPUs with no communication)'.format(args.gpus),
              file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.cpu_baseline_child:
        print(json.dumps(cpu_baseline_child(args.size, args.emitters)))
        return
    import numpy
    import torch
    import scopyon_amd
    from scopyon_amd.batched import BatchedSimulator

    dt = 0.033
    cfg = _config(args.size, dt)
    F, S = args.fovs, args.size
    sim = BatchedSimulator(cfg, [(S, S)] * F, args.emitters, D=0.1e-12, dt=dt, seed=1234,
                           boundary='reflect')
    sim.capture()
    for _ in range(args.warmup):
        sim.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.replay(args.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    assert sim.status() & ~1 == 0
    batched_fps = F * args.steps / elapsed

    legs = set() if args.legs == 'none' else set(args.legs.split(','))
    plain_fps = api_fps = None
    if 'plain' in legs:
        # plain single-frame path: same kernels, one frame per step, eager launches
        plain = BatchedSimulator(cfg, [(S, S)], args.emitters, D=0.1e-12, dt=dt, seed=99,
                                 boundary='reflect')
        for _ in range(3):
            plain.step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.plain_steps):
            plain.step()
        torch.cuda.synchronize()
        plain_fps = args.plain_steps / (time.perf_counter() - t0)
        del plain
    if 'api' in legs:
        # drop-in API: form_image with a RandomState (reference-exact EMCCD draws)
        rng = numpy.random.RandomState(5)
        pl = 16e-6 / 100.0
        inputs = rng.uniform(-0.5 * S * pl, 0.5 * S * pl, size=(args.emitters, 2))
        scopyon_amd.form_image(inputs, config=cfg, rng=rng)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        napi = 3
        for _ in range(napi):
            scopyon_amd.form_image(inputs, config=cfg, rng=rng)
        api_fps = napi / (time.perf_counter() - t0)

    # per-stage launch times (graph-replayed back-to-back launches, HIP events
    # on the launch stream) and the dominant kernel's roofline
    kt = _kernel_times(sim, 20)
    sbytes = _stage_bytes(sim.n, sim.n_pix)
    dom = max(kt, key=kt.get)
    conv, impl = sbytes[dom]
    sec = kt[dom] * 1e-3
    ms_step = elapsed / args.steps * 1e3
    result = {
        'metric': METRIC,
        'value': batched_fps,
        'unit': 'frames/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': ms_step,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32 table / f64 accumulate / f32 sampler',
        'data': 'synthetic',
        'config': {'workload': 'batched TIRF EMCCD frames, {0}x{0} px, {1} emitters per FOV, '
                               '{2} FOVs per step, hipGraph replay'.format(S, args.emitters, F),
                   'fovs': F, 'frame': [S, S], 'emitters_per_fov': args.emitters,
                   'pixel_nm': 160,
                   'psf': 'TRITC Born-Wolf; folded float32 phase table from the float64 SAT; '
                          'float64 accumulation',
                   'dt_s': dt, 'D_um2_s': 0.1, 'walls': 'reflecting',
                   'emccd': 'Philox4x32-10, float32 sampler' +
                            (', fused into the splat tile pass' if sim.fused else ''),
                   'tile_side': sim.tile_side, 'quadrant': sim.quadrant,
                   'tile_autotune_ms': {str(k): v for k, v in sim.tile_timings.items()}},
        # plain = the same kernels, one frame per step, eager launches
        'plain_frames_per_sec': plain_fps,
        'batched_over_plain': batched_fps / plain_fps if plain_fps else None,
        # drop-in form_image(): RandomState draws on the host, reference-exact EMCCD
        'api_form_image_frames_per_sec': api_fps,
        'batched_over_api_form_image': batched_fps / api_fps if api_fps else None,
        'kernel_ms': kt,
        'kernel_ms_sum': sum(kt.values()),
        'roofline': {'bound': 'hbm', 'kernel': KERNELS.get(dom, dom), 'stage': dom,
                     'achieved': conv / sec / 1e9, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': conv / sec / 1e9 / HBM_PEAK_GBS, 'traffic': _traffic(dom),
                     'algorithmic_bytes': conv,
                     'algorithmic_bytes_note': '16 N + 4 HW per frame (emitter data in, '
                                               'float32 ADC out)',
                     'implementation_bytes': impl,
                     'frac_implementation_bytes': impl / sec / 1e9 / HBM_PEAK_GBS,
                     'launch_ms': kt[dom]},
    }
    _progress('headline: {:.0f} frames/s, {:.3f} ms per step; stage times {}'.format(
        batched_fps, ms_step, json.dumps(kt)))
    if 'uneven' in legs:
        result['batched_uneven'] = batched_uneven_leg()
        _progress('batched_uneven done')
    if 'scale3d' in legs:
        result['scale_3d'] = scale3d_leg()
        _progress('scale_3d done')
    if 'precision' in legs:
        result['precision'] = precision_leg()
        _progress('precision done')
    if 'fwd_bwd' in legs:
        result['fwd_bwd'] = fwd_bwd()
        result['fwd_bwd_per_sec'] = result['fwd_bwd']['fwd_bwd_per_sec']
        _progress('fwd_bwd done: {:.0f} iterations/s'.format(result['fwd_bwd_per_sec']))
    result['max_rel_err_vs_fp64_oracle'] = gpu_vs_oracle_error(S)
    _progress('error vs the float64 oracle: {:.2e}'.format(result['max_rel_err_vs_fp64_oracle']))
    if not args.no_cpu:
        base = cpu_baseline(S, args.emitters)
        result['cpu_baseline'] = base
        result['gpu_over_cpu'] = batched_fps / base['value']
        _progress('cpu baseline done')
    print(json.dumps(result))


if __name__ == '__main__':
    main()
