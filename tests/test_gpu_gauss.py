"""Differentiable Gaussian model on the GPU against the float64 oracle:
forward image, Poisson NLL, hand-written backward, float64 central finite
differences, autograd plumbing, and an Adam refinement of positions."""
import numpy
import pytest
import torch

from oracle import gauss as G

pytestmark = pytest.mark.gpu

KW = dict(sigma=1.2, inv_depth=1.0 / 0.377, background=5.0, radius=6)


def _case(seed, shapes, n_per, shift=(0.25, -0.15)):
    rng = numpy.random.RandomState(seed)
    xs, ys, fid = [], [], []
    for f, (nw, nh) in enumerate(shapes):
        xs.append(rng.uniform(-2.0, nw + 2.0, n_per))
        ys.append(rng.uniform(-2.0, nh + 2.0, n_per))
        fid.append(numpy.full(n_per, f))
    x, y, fid = numpy.concatenate(xs), numpy.concatenate(ys), numpy.concatenate(fid)
    n = len(x)
    z = rng.uniform(-0.5, 0.5, n)
    I = rng.uniform(300.0, 900.0, n)
    truth = G.render(shapes, x + shift[0], y + shift[1], z, I, fid, **KW)
    k = [rng.poisson(m).astype(float) for m in truth]
    return x, y, z, I, fid, k


def _dev(a, dtype=torch.float32):
    return torch.as_tensor(numpy.ascontiguousarray(a), dtype=dtype, device='cuda')


def _model(shapes, method='mfma-f32'):
    from scopyon_amd.gauss import GaussianModel
    return GaussianModel(shapes, method=method, **KW)


@pytest.mark.parametrize('method', ['mfma-f32', 'lds'])
def test_forward_and_nll_match_oracle(gpu, method):
    shapes = [(64, 64), (40, 72), (100, 33)]
    x, y, z, I, fid, k = _case(0, shapes, 60)
    m = _model(shapes, method)
    mu = m.forward(_dev(x), _dev(y), _dev(z), _dev(I), _dev(fid, torch.int32))
    assert m.check_status() & ~1 == 0
    ref = G.render(shapes, x, y, z, I, fid, **KW)
    for f in range(len(shapes)):
        got = m.frame(mu, f).cpu().numpy().astype(float)
        numpy.testing.assert_allclose(got, ref[f], rtol=2e-6, atol=1e-6 * ref[f].max())
    kd = torch.cat([_dev(a.ravel()) for a in k])
    grad = torch.empty_like(mu)
    loss = m.nll(mu, kd, grad).cpu().numpy()
    want = G.poisson_nll([m.frame(mu, f).cpu().numpy().astype(float) for f in range(3)],
                         [a.astype(numpy.float32).astype(float) for a in k])
    numpy.testing.assert_allclose(loss, want, rtol=1e-12)
    mu_h = mu.cpu().numpy().astype(numpy.float32)
    # 1 - k * rcp(mu) on the hardware reciprocal (1 ulp): within 1e-6
    # relative plus two float32 ulps of 1 where 1 - k/mu cancels
    numpy.testing.assert_allclose(grad.cpu().numpy(), 1.0 - kd.cpu().numpy() / mu_h, rtol=1e-6,
                                  atol=2.5e-7)


@pytest.mark.parametrize('method', ['mfma-f32', 'mfma-bf16'])
def test_forward_nll_fused_matches_separate(gpu, method):
    """Forward + NLL in one matrix-core pass: the gradient 1 - k/mu is
    bit-identical to forward() then nll(), with and without writing the
    image; the per-frame loss (terms with the float32 hardware logarithm,
    summed in float64 per tile, then over tiles) within 2e-6 of nll()'s
    float64-log loss."""
    shapes = [(64, 64), (40, 72), (100, 33)]
    x, y, z, I, fid, k = _case(0, shapes, 60)
    m = _model(shapes, method)
    args = [_dev(x), _dev(y), _dev(z), _dev(I)]
    kd = torch.cat([_dev(a.ravel()) for a in k])
    mu = m.forward(*args, _dev(fid, torch.int32))
    grad = torch.empty_like(mu)
    want = m.nll(mu, kd, grad).clone()
    for write_mu in (True, False):
        mu2 = torch.full_like(mu, -1.0) if write_mu else None
        grad2 = torch.full_like(mu, -7.0)
        loss = m.forward_nll(*args, kd, _dev(fid, torch.int32), grad_mu=grad2, mu=mu2)
        assert m.check_status() & ~1 == 0
        torch.testing.assert_close(loss, want, rtol=2e-6, atol=0)
        assert torch.equal(grad2, grad)
        if write_mu:
            assert torch.equal(mu2, mu)


@pytest.mark.parametrize('method', ['lds', 'mfma-f32'])
def test_many_frames_span_the_binning_window(gpu, method):
    """72 frames of 256^2 (64 tiles each, 4608 in all) with 16 emitters each:
    a 1024-emitter binning block spans more tiles than its LDS counting window
    (4096), so tile counts and slot claims mix the LDS-aggregated and the
    direct global path; every frame still equals the oracle."""
    shapes = [(256, 256)] * 72
    x, y, z, I, fid, _ = _case(5, shapes, 16)
    m = _model(shapes)
    assert m.n_tiles > 4096 and len(x) > 1024
    zt = _dev(z) if method == 'lds' else None
    mu = m.forward(_dev(x), _dev(y), zt, _dev(I), _dev(fid, torch.int32), method=method)
    assert m.check_status() & ~1 == 0
    zr = z if method == 'lds' else numpy.zeros_like(z)
    ref = G.render(shapes, x, y, zr, I, fid, **KW)
    for f in range(len(shapes)):
        got = m.frame(mu, f).cpu().numpy().astype(float)
        numpy.testing.assert_allclose(got, ref[f], rtol=2e-5, atol=1e-5 * ref[f].max(),
                                      err_msg=str(f))


@pytest.mark.parametrize('method', ['mfma-f32', 'lds'])
def test_forward_is_deterministic(gpu, method):
    shapes = [(96, 96)]
    x, y, z, I, fid, _ = _case(1, shapes, 400)
    m = _model(shapes, method)
    args = [_dev(x), _dev(y), _dev(z), _dev(I), None]
    a = m.forward(*args).clone()
    b = m.forward(*args).clone()
    assert torch.equal(a, b)


def _gpu_grads(m, x, y, z, I, fid, k):
    mu = m.forward(_dev(x), _dev(y), _dev(z), _dev(I), _dev(fid, torch.int32))
    kd = torch.cat([_dev(a.ravel()) for a in k])
    grad = torch.empty_like(mu)
    m.nll(mu, kd, grad)
    gx, gy, gz, gI, gs = m.backward(_dev(x), _dev(y), _dev(z), _dev(I), grad,
                                    _dev(fid, torch.int32))
    return [t.cpu().numpy().astype(float) for t in (gx, gy, gz, gI)] + [float(gs.item())]


@pytest.mark.parametrize('method', ['mfma-f32', 'lds', 'fused-nll'])
def test_binned_backward_equals_standalone(gpu, method):
    """The binned backward (each tile reduces the emitters whose clamped
    window centre it holds, four lanes per emitter, reading its LDS window
    of dL/dmu, from the lists the forward sorted) against the standalone
    backward (16 lanes per emitter, another float32 summation order):
    per-emitter gradients within 1e-5 relative (plus 1e-6 of the largest),
    every output written -- emitters whose window crosses a frame edge, lies
    off every frame, with a frame id out of range or a NaN position included
    -- bit-repeatable, and dL/dsigma within 1e-6.  A binning of another sigma
    raises status bit 16."""
    from scopyon_amd import _device, _lib
    import ctypes
    shapes = [(64, 64), (40, 72), (100, 33)]
    x, y, z, I, fid, k = _case(3, shapes, 80)
    ex = numpy.array([-20.0, -3.5, 44.2, 10.0, 10.0, numpy.nan, 63.9, -0.5])
    ey = numpy.array([5.0, 30.0, 71.5, 10.0, 10.0, 3.0, -6.9, 32.99])
    ef = numpy.array([0, 0, 1, 7, -1, 2, 0, 2])
    x, y, fid = numpy.concatenate([x, ex]), numpy.concatenate([y, ey]), numpy.concatenate([fid, ef])
    z = numpy.concatenate([z, numpy.full(len(ex), 0.1)])
    I = numpy.concatenate([I, numpy.full(len(ex), 500.0)])
    n = len(x)
    m = _model(shapes, 'lds' if method == 'lds' else 'mfma-f32')
    X = [_dev(a) for a in (x, y, z, I)]
    F = _dev(fid, torch.int32)
    kd = torch.cat([_dev(a.ravel()) for a in k])
    grad = torch.empty(m.n_pix, dtype=torch.float32, device='cuda')
    if method == 'fused-nll':
        m.forward_nll(*X, kd, F, grad_mu=grad)
    else:
        mu = m.forward(*X, F)
        m.nll(mu, kd, grad)

    def run(binned):
        outs = [torch.full((n,), float('nan'), device='cuda') for _ in range(4)]
        r = m.backward(*X, grad, F, outputs=outs, binned=binned)
        return [t.clone() for t in r[:4]], float(r[4].item())
    gb, sb = run(True)
    gb2, sb2 = run(True)
    assert m.check_status() & ~1 == 0
    gs, ss = run(False)
    for a, b, c, name in zip(gb, gb2, gs, 'x y z I'.split()):
        assert not torch.isnan(a).any(), name
        assert torch.equal(a, b), name
        an, cn = a.cpu().numpy().astype(float), c.cpu().numpy().astype(float)
        numpy.testing.assert_allclose(an, cn, rtol=1e-5, atol=1e-6 * numpy.abs(cn).max(),
                                      err_msg=name)
    assert sb == sb2
    assert abs(sb - ss) <= 1e-6 * abs(ss)
    assert (gs[3][-8:].cpu().numpy()[[0, 3, 4, 5]] == 0).all()
    # the standalone backward reused the workspace: no binning left to reuse
    with pytest.raises(RuntimeError):
        m.backward(*X, grad, F, binned=True)
    m.forward(*X, F)
    p = m.params(1.3)
    outs = [torch.empty(n, device='cuda') for _ in range(4)]
    ws, cap = m._workspace(n)
    _lib.call('smx_gauss_backward_binned', n, X[0].data_ptr(), X[1].data_ptr(), X[2].data_ptr(),
              X[3].data_ptr(), F.data_ptr(), m.frames_dev.data_ptr(), m.nframes, m.n_tiles,
              ctypes.byref(p), cap, grad.data_ptr(), *[o.data_ptr() for o in outs],
              m.gsigma.data_ptr(), ws.data_ptr(), ws.numel(), m.status.data_ptr(),
              _device.stream())
    assert int(m.status.item()) & 16
    m.status.zero_()


def test_backward_matches_oracle_gradients(gpu):
    shapes = [(48, 48), (32, 60)]
    x, y, z, I, fid, k = _case(2, shapes, 40)
    m = _model(shapes)
    got = _gpu_grads(m, x, y, z, I, fid, k)
    kf = [a.astype(numpy.float32).astype(float) for a in k]
    want = G.gradients(shapes, x, y, z, I, fid, targets=kf, **KW)
    for g, w, name in zip(got[:4], want[:4], 'x y z I'.split()):
        scale = numpy.abs(w).max()
        numpy.testing.assert_allclose(g, w, rtol=1e-4, atol=1e-4 * scale, err_msg=name)
    assert abs(got[4] - want[4]) <= 1e-4 * abs(want[4]) + 1e-3


def test_backward_against_float64_central_differences(gpu):
    """GPU float32 gradients vs float64 central differences of the oracle
    (h = 1e-4 px for positions, 1e-4 um for z, 1e-2 for I, 1e-4 px for sigma),
    both at the float32 parameters the GPU sees.  Error bound: 1e-4 relative,
    measured against max(|fd|, RMS of that parameter's gradients) -- a float32
    image has a rounding floor of ~1e-7 per pixel, so components that nearly
    cancel cannot be resolved to 1e-4 of their own size."""
    shapes = [(40, 40)]
    x, y, z, I, fid, k = _case(3, shapes, 25)
    x, y, z, I = (a.astype(numpy.float32).astype(float) for a in (x, y, z, I))
    sig32 = float(numpy.float32(KW['sigma']))
    m = _model(shapes)
    got = _gpu_grads(m, x, y, z, I, fid, k)
    kf = [a.astype(numpy.float32).astype(float) for a in k]
    base = dict(sigma=sig32, inv_depth=float(numpy.float32(KW['inv_depth'])),
                background=KW['background'], radius=KW['radius'])

    def L(xx, yy, zz, II, sig):
        b = dict(base)
        b['sigma'] = sig
        return G.loss(shapes, xx, yy, zz, II, fid, targets=kf, **b)

    for idx, h in enumerate((1e-4, 1e-4, 1e-4, 1e-2)):
        fds, gs = [], []
        for e in range(len(x)):
            p, q = [a.copy() for a in (x, y, z, I)], [a.copy() for a in (x, y, z, I)]
            p[idx][e] += h
            q[idx][e] -= h
            fds.append((L(*p, sig32) - L(*q, sig32)) / (2 * h))
            gs.append(got[idx][e])
        fds, gs = numpy.array(fds), numpy.array(gs)
        rms = numpy.sqrt(numpy.mean(fds ** 2))
        rel = numpy.abs(gs - fds) / numpy.maximum(numpy.abs(fds), rms)
        assert rel.max() <= 1e-4, ('xyzI'[idx], rel.max(), int(rel.argmax()))
        big = numpy.abs(fds) >= 0.1 * numpy.abs(fds).max()
        assert (numpy.abs(gs - fds)[big] / numpy.abs(fds[big])).max() <= 1e-4
    fd = (L(x, y, z, I, sig32 + 1e-4) - L(x, y, z, I, sig32 - 1e-4)) / 2e-4
    assert abs(got[4] - fd) <= 1e-4 * abs(fd), ('sigma', got[4], fd)


def test_autograd_function_and_adam_refinement(gpu):
    from scopyon_amd.gauss import render, poisson_nll
    shapes = [(64, 64), (64, 64)]
    rng = numpy.random.RandomState(4)
    n_per = 30
    tx = rng.uniform(8, 56, 2 * n_per)
    ty = rng.uniform(8, 56, 2 * n_per)
    fid = numpy.repeat([0, 1], n_per)
    z = rng.uniform(-0.3, 0.3, 2 * n_per)
    I = numpy.full(2 * n_per, 2000.0)
    truth = G.render(shapes, tx, ty, z, I, fid, **KW)
    k = torch.cat([_dev(rng.poisson(t).astype(float).ravel()) for t in truth])
    m = _model(shapes)
    x = _dev(tx + rng.normal(0, 0.4, 2 * n_per)).requires_grad_(True)
    y = _dev(ty + rng.normal(0, 0.4, 2 * n_per)).requires_grad_(True)
    zt, It, f32 = _dev(z), _dev(I), _dev(fid, torch.int32)
    # the autograd path equals the raw kernels
    mu = render(m, x, y, zt, It, None, f32)
    loss = poisson_nll(m, mu, k).sum()
    loss.backward()
    mu2 = m.forward(x.detach(), y.detach(), zt, It, f32)
    g2 = torch.empty_like(mu2)
    m.nll(mu2, k, g2)
    gx, gy, _, _, _ = m.backward(x.detach(), y.detach(), zt, It, g2, f32)
    assert torch.allclose(x.grad, gx) and torch.allclose(y.grad, gy)
    err0 = float(numpy.sqrt(numpy.mean((x.detach().cpu().numpy() - tx) ** 2 +
                                       (y.detach().cpu().numpy() - ty) ** 2)))
    opt = torch.optim.Adam([x, y], lr=0.05)
    first = None
    for it in range(200):
        opt.zero_grad()
        loss = poisson_nll(m, render(m, x, y, zt, It, None, f32), k).sum()
        loss.backward()
        opt.step()
        first = float(loss.detach()) if first is None else first
    err = float(numpy.sqrt(numpy.mean((x.detach().cpu().numpy() - tx) ** 2 +
                                      (y.detach().cpu().numpy() - ty) ** 2)))
    assert float(loss.detach()) < first
    assert err < 0.25 * err0, (err0, err)


def test_sigma_gradient_through_autograd(gpu):
    from scopyon_amd.gauss import render, poisson_nll
    shapes = [(48, 48)]
    x, y, z, I, fid, k = _case(5, shapes, 30)
    m = _model(shapes)
    sig = torch.tensor(KW['sigma'], dtype=torch.float32, device='cuda', requires_grad=True)
    kd = _dev(k[0].ravel())
    loss = poisson_nll(m, render(m, _dev(x), _dev(y), _dev(z), _dev(I), sig, None), kd).sum()
    loss.backward()
    kf = [k[0].astype(numpy.float32).astype(float)]
    want = G.gradients(shapes, x, y, z, I, fid, targets=kf, **KW)[4]
    assert abs(float(sig.grad) - want) <= 1e-4 * abs(want) + 1e-3


def test_bad_arguments(gpu):
    from scopyon_amd.gauss import GaussianModel
    with pytest.raises(ValueError):
        GaussianModel([(8, 8)], radius=9)
    m = GaussianModel([(8, 8), (8, 8)])
    with pytest.raises(ValueError):
        m.forward(_dev([1.0]), _dev([1.0]), _dev([0.0]), _dev([1.0]))
    with pytest.raises(TypeError):
        m.forward(_dev([1.0], torch.float64), _dev([1.0]), _dev([0.0]), _dev([1.0]),
                  _dev([0], torch.int32))


@pytest.mark.parametrize('method,rms_tol,max_tol', [('mfma-f32', 2e-6, 2e-5),
                                                    ('mfma-f16', 1e-3, 5e-3),
                                                    ('mfma-bf16', 8e-3, 4e-2),
                                                    ('lds', 2e-6, 2e-5)])
def test_separable_forms_against_oracle(gpu, method, rms_tol, max_tol):
    """Matrix-core contraction (float32/float16/bfloat16 operands, float32
    accumulation) and the LDS splat vs the float64 oracle: relative RMS error
    ||mu - ref|| / ||ref|| and the largest relative error over pixels above 1%
    of the frame maximum (intensities exponential, mean 500; no background)."""
    from scopyon_amd.gauss import GaussianModel
    shapes = [(128, 96), (64, 64)]
    rng = numpy.random.RandomState(6)
    n_per = 300
    x = numpy.concatenate([rng.uniform(-3, s[0] + 3, n_per) for s in shapes])
    y = numpy.concatenate([rng.uniform(-3, s[1] + 3, n_per) for s in shapes])
    fid = numpy.repeat(numpy.arange(len(shapes)), n_per)
    I = rng.exponential(500.0, len(x))
    z = numpy.zeros(len(x))
    kw = dict(sigma=1.2, inv_depth=0.0, background=0.0, radius=6)
    m = GaussianModel(shapes, method=method, **kw)
    zt = None if method != 'lds' else _dev(z)
    mu = m.forward(_dev(x), _dev(y), zt, _dev(I), _dev(fid, torch.int32))
    assert m.check_status() & ~1 == 0
    ref = G.render(shapes, x, y, z, I, fid, **kw)
    for f in range(len(shapes)):
        got = m.frame(mu, f).cpu().numpy().astype(float)
        rms = numpy.linalg.norm(got - ref[f]) / numpy.linalg.norm(ref[f])
        big = ref[f] > 0.01 * ref[f].max()
        mx = (numpy.abs(got - ref[f])[big] / ref[f][big]).max()
        assert rms <= rms_tol and mx <= max_tol, (method, f, rms, mx)


@pytest.mark.parametrize('method', ['lds', 'mfma-f32'])
def test_large_launch_binning_against_oracle(gpu, method):
    """600,000 emitters on one 1024^2 frame: the binning takes four emitters
    per thread (launches of 2^19 emitters and more) and tile lists of ~1,300
    entries; a 48 x 48 crop equals the oracle (computed from the emitters whose
    window reaches it) and the frame is deterministic."""
    from scopyon_amd.gauss import GaussianModel
    S, N, R, C = 1024, 600_000, 6, 48
    rng = numpy.random.RandomState(8)
    x, y = rng.uniform(0, S, N), rng.uniform(0, S, N)
    I = rng.exponential(500.0, N)
    kw = dict(sigma=1.2, inv_depth=0.0, background=0.0, radius=R)
    m = GaussianModel([(S, S)], method=method, **kw)
    xd, yd, Id = _dev(x), _dev(y), _dev(I)
    zt = _dev(numpy.zeros(N)) if method == 'lds' else None
    mu = m.forward(xd, yd, zt, Id)
    mu2 = m.forward(xd, yd, zt, Id).clone()
    assert m.check_status() & ~1 == 0
    assert torch.equal(mu, mu2)
    x32, y32, I32 = (a.astype(numpy.float32).astype(float) for a in (x, y, I))
    c0 = 500
    sel = ((x32 >= c0 - R - 1) & (x32 < c0 + C + R + 1) & (y32 >= c0 - R - 1) &
           (y32 < c0 + C + R + 1))
    k = int(sel.sum())
    ref = G.render([(C, C)], x32[sel] - c0, y32[sel] - c0, numpy.zeros(k), I32[sel],
                   numpy.zeros(k, int), **kw)[0]
    got = mu.view(S, S)[c0:c0 + C, c0:c0 + C].cpu().numpy().astype(float)
    rms = numpy.linalg.norm(got - ref) / numpy.linalg.norm(ref)
    assert rms <= 2e-6, (method, rms)


def test_mfma_forward_deterministic_and_unsorted_lists(gpu):
    """Repeatable bits for sorted tile lists, also for lists longer than the
    tile pass sorts in LDS (2,048 entries: sorted after binning); a list of
    more than 8,192 entries keeps the image (summation order aside) and
    reports status bit 1."""
    from scopyon_amd.gauss import GaussianModel
    shapes = [(64, 64)]
    rng = numpy.random.RandomState(7)
    kw = dict(sigma=1.2, inv_depth=0.0, background=1.0, radius=6)
    x, y = rng.uniform(0, 64, 600), rng.uniform(0, 64, 600)
    for extra in (0, 2500, 9000):
        print('gauss lists: {} extra emitters'.format(extra), flush=True)
        xs = numpy.concatenate([rng.uniform(8, 24, extra), x])
        ys = numpy.concatenate([rng.uniform(8, 24, extra), y])
        I = rng.uniform(1.0, 2.0, len(xs))
        m = GaussianModel(shapes, method='mfma-f32', **kw)
        args = (_dev(xs), _dev(ys), None, _dev(I))
        a = m.forward(*args).clone()
        if extra <= 2500:
            b = m.forward(*args).clone()
            assert torch.equal(a, b)
            assert m.check_status() == 0
        else:
            assert m.check_status() & 1
        ref = G.render(shapes, xs, ys, numpy.zeros(len(xs)), I, numpy.zeros(len(xs), int), **kw)[0]
        numpy.testing.assert_allclose(a.view(64, 64).cpu().numpy(), ref, rtol=2e-5)


def test_mfma_forward_merge_sorted_lists(gpu):
    """2,000 emitters on 64^2 (four 32 px tiles): ~700 entries per tile list,
    sorted in the tile pass by the merge-rank sort's four 256-entry runs --
    repeatable bits and the image of the float64 renderer."""
    from scopyon_amd.gauss import GaussianModel
    shapes = [(64, 64)]
    rng = numpy.random.RandomState(29)
    kw = dict(sigma=1.2, inv_depth=0.0, background=1.0, radius=6)
    n = 2000
    x, y = rng.uniform(0, 64, n), rng.uniform(0, 64, n)
    I = rng.uniform(1.0, 2.0, n)
    m = GaussianModel(shapes, method='mfma-f32', **kw)
    args = (_dev(x), _dev(y), None, _dev(I))
    a = m.forward(*args).clone()
    b = m.forward(*args).clone()
    assert torch.equal(a, b)
    assert m.check_status() == 0
    ref = G.render(shapes, x, y, numpy.zeros(n), I, numpy.zeros(n, int), **kw)[0]
    numpy.testing.assert_allclose(a.view(64, 64).cpu().numpy(), ref, rtol=2e-5)
