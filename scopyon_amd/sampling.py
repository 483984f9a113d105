"""Brownian input sampling, legacy API (src/scopyon/sampling.py:14-162).

With a ``numpy.random.RandomState`` every draw is taken from it in the
reference's order (uniform positions dim by dim, then one normal per particle
and axis per time step) and the trajectory arithmetic runs on the GPU, giving
bit-identical results.  With ``PhiloxRNG`` everything is drawn on the device.
"""
import collections.abc
import numbers
import warnings

import numpy
import torch

from . import _device, _lib
from .rng import PhiloxRNG

__all__ = ["sample_inputs"]


def _limits(lower, upper, ndim):
    lower = (numpy.ones(ndim) * lower if isinstance(lower, numbers.Number)
             else numpy.array(lower, dtype=float) if lower is not None else numpy.zeros(ndim))
    upper = (numpy.ones(ndim) * upper if isinstance(upper, numbers.Number)
             else numpy.array(upper, dtype=float) if upper is not None else numpy.ones(ndim))
    if len(lower) < ndim or len(upper) < ndim:
        raise ValueError("The wrong size of limits was given [(lower={}, upper={}) != {}].".format(
            len(lower), len(upper), ndim))
    for d in range(ndim):
        if lower[d] > upper[d]:
            lower[d], upper[d] = upper[d], lower[d]
    return lower, upper


def sample_points(rng, N=None, conc=None, lower=None, upper=None, start=0, ndim=3):
    """sampling.py:14-86 -> (points [coords, id, p_state=1], last id)."""
    if N is None and conc is None:
        raise ValueError('Either one of N or conc must be given.')
    lower, upper = _limits(lower, upper, ndim)
    if N is None:
        lengths = upper - lower
        size = numpy.prod(lengths[lengths != 0])
        concs = conc if isinstance(conc, collections.abc.Iterable) else [conc]
        if isinstance(rng, PhiloxRNG):
            counts = [int(rng._host.poisson(size * c)) for c in concs]
        else:
            counts = [rng.poisson(size * c) for c in concs]
    elif not isinstance(N, collections.abc.Iterable):
        counts = [N]
    else:
        counts = N
    N = int(sum(counts))
    if N <= 0:
        return numpy.array([]), start
    maxdim = len(lower)
    ret = numpy.zeros((N, maxdim + 2))
    if isinstance(rng, PhiloxRNG):
        dev = _device.require_gpu()
        xs = [torch.empty(N, dtype=torch.float64, device=dev) for _ in range(3)]
        if maxdim > 3:
            raise NotImplementedError('PhiloxRNG sampling supports ndim <= 3')
        lo = _lib.host_doubles(lower[:3], 3)
        hi = _lib.host_doubles(upper[:3], 3)
        _lib.call('smx_uniform_init', N, maxdim, xs[0].data_ptr(), xs[1].data_ptr(),
                  xs[2].data_ptr(), lo, hi, rng.seed, rng.next_counter(), 0, _device.stream())
        for d in range(min(maxdim, 3)):
            ret[:, d] = xs[d].cpu().numpy()
    else:
        for d in range(maxdim):
            if lower[d] < upper[d]:
                ret[:, d] = rng.uniform(lower[d], upper[d], N)
            else:
                ret[:, d] = lower[d]
    ret[:, maxdim] = numpy.arange(start, start + N)
    ret[:, maxdim + 1] = 1.0
    return ret, start + N


def _axis_D(D, ndim):
    if D is None:
        return numpy.zeros(ndim)
    if not isinstance(D, collections.abc.Iterable):
        return numpy.ones(ndim) * D
    D = numpy.asarray(D)
    assert len(D) == ndim
    return D


def trajectory_predrawn(x0, scale, xi, boundary=0, lower=None, upper=None, state0=None,
                        pacc=None, trans_u=None):
    """Run smx_brownian_predrawn: x0 (n, ndim), scale (T, nstates, ndim), xi (T, n, ndim).
    Returns positions (T+1, n, ndim) and states (T+1, n)."""
    dev = _device.require_gpu()
    n, ndim = x0.shape
    T = scale.shape[0]
    nstates = scale.shape[1] if T else (1 if pacc is None else pacc.shape[1])
    xs = torch.empty((T + 1, n, ndim), dtype=torch.float64, device=dev)
    states = torch.empty((T + 1, n), dtype=torch.int32, device=dev)
    d_x0 = _device.upload(x0, device=dev)
    d_scale = _device.upload(scale if T else numpy.zeros((1, nstates, ndim)), device=dev)
    d_xi = _device.upload(xi if T else numpy.zeros((1, n, ndim)), device=dev)
    d_s0 = None if state0 is None else _device.upload(state0, dtype=torch.int32, device=dev)
    d_lo = None if lower is None else _device.upload(lower, device=dev)
    d_hi = None if upper is None else _device.upload(upper, device=dev)
    d_pacc = None if pacc is None else _device.upload(pacc, device=dev)
    d_tu = None if trans_u is None else _device.upload(trans_u, device=dev)
    _lib.call('smx_brownian_predrawn', n, ndim, T, nstates, d_x0.data_ptr(), _device.ptr(d_s0),
              d_scale.data_ptr(), d_xi.data_ptr(), boundary, _device.ptr(d_lo), _device.ptr(d_hi),
              _device.ptr(d_pacc), _device.ptr(d_tu), xs.data_ptr(), states.data_ptr(),
              _device.stream())
    return xs.cpu().numpy(), states.cpu().numpy()


def sample_inputs(t, *, N=None, conc=None, lower=None, upper=None, D=None, start=0, ndim=3,
                  rng=None):
    """sampling.py:121-162 -> [(t_k, (N, ndim + 2) points)]."""
    if rng is None:
        warnings.warn('A random number generator [rng] is not given.')
        rng = numpy.random.RandomState()
    t = sorted(t)
    points, _ = sample_points(rng, N=N, conc=conc, lower=lower, upper=upper, start=start,
                              ndim=ndim)
    if points.size == 0:
        return [(t[0] if i == 0 else max(t[:i + 1]), points.copy()) for i in range(len(t))]
    Dax = _axis_D(D, ndim)
    n = points.shape[0]
    # time points that move the particles (tnext > tcurrent)
    moves, tcur = [], t[0]
    for tnext in t:
        if tnext > tcur:
            moves.append(tnext - tcur)
            tcur = tnext
    T = len(moves)
    x0 = numpy.ascontiguousarray(points[:, :ndim])
    if isinstance(rng, PhiloxRNG):
        traj, _ = _philox_walk(rng, x0, numpy.asarray(Dax, dtype=float)[None, :], moves)
    else:
        scale = numpy.empty((T, 1, ndim))
        xi = numpy.empty((T, n, ndim))
        for k, dt in enumerate(moves):
            scale[k, 0, :] = [numpy.sqrt(2 * Dax[d] * dt) for d in range(ndim)]
            xi[k] = rng.standard_normal(n * ndim).reshape(n, ndim)
        traj, _ = trajectory_predrawn(x0, scale, xi)
    out, k, tcur = [], 0, t[0]
    for tnext in t:
        if tnext > tcur:
            k += 1
            tcur = tnext
        pts = points.copy()
        pts[:, :ndim] = traj[k]
        out.append((tcur, pts))
    return out


def _philox_walk(rng, x0, D_tab, moves, boundary=0, lower=None, upper=None, state0=None,
                 transmat=None):
    """Device Philox trajectory: x0 (n, ndim <= 3), D_tab (nstates, ndim)
    diffusion per state and axis, state0 (n,) initial states, transmat
    (nstates, nstates) rates [1/s] (zero diagonal) for state transitions
    after each move (sampling2.py:52-68).  Returns (trajectory (T+1, n, ndim),
    states (T+1, n))."""
    dev = _device.require_gpu()
    n, ndim = x0.shape
    if ndim > 3:
        raise NotImplementedError('PhiloxRNG sampling supports ndim <= 3')
    D_tab = numpy.asarray(D_tab, dtype=float).reshape(-1, ndim)
    nstates = D_tab.shape[0]
    xs = [_device.upload(x0[:, d].copy(), device=dev) if d < ndim else
          torch.zeros(1, dtype=torch.float64, device=dev) for d in range(3)]
    traj = numpy.empty((len(moves) + 1, n, ndim))
    traj[0] = x0
    s0 = numpy.zeros(n, dtype=numpy.int32) if state0 is None else numpy.asarray(state0, numpy.int32)
    states = numpy.empty((len(moves) + 1, n), dtype=numpy.int32)
    states[0] = s0
    d_state = _device.upload(s0, dtype=torch.int32, device=dev)
    lo = _lib.host_doubles(lower[:ndim] if lower is not None else [], 3)
    hi = _lib.host_doubles(upper[:ndim] if upper is not None else [], 3)
    d_D = _device.upload(numpy.ascontiguousarray(D_tab), device=dev)
    for k, dt in enumerate(moves):
        d_pacc = None
        if transmat is not None:
            P = 1 - numpy.exp(-transmat * dt)
            assert (P.sum(axis=1) <= 1.0).all()
            P.ravel()[::nstates + 1] = 1.0 - P.sum(axis=1)
            d_pacc = _device.upload(numpy.ascontiguousarray(P.cumsum(axis=1)), device=dev)
        _lib.call('smx_brownian_philox_states', n, ndim, xs[0].data_ptr(), xs[1].data_ptr(),
                  xs[2].data_ptr(), d_state.data_ptr(), nstates, d_D.data_ptr(), float(dt),
                  boundary, lo, hi, _device.ptr(d_pacc), rng.seed, rng.next_counter(), 0,
                  _device.stream())
        for d in range(ndim):
            traj[k + 1, :, d] = xs[d].cpu().numpy()
        states[k + 1] = d_state.cpu().numpy()
    return traj, states
