"""Brownian sampling with states (src/scopyon/sampling2.py:15-149).

``sample`` keeps the reference's signature and output ([coords, state, id]
per time point).  Extension: ``boundary='reflect'`` (reflecting walls on
[lower, upper]); ``periodic=True`` is the reference's periodic wrap.
"""
import collections.abc
import warnings

import numpy

from .rng import PhiloxRNG
from .sampling import _limits, _philox_walk, trajectory_predrawn

__all__ = ["sample"]


def _generate_points(rng, N, lower, upper, ndim):
    """sampling2.py:15-31 (uniforms drawn group -> dim -> particle)."""
    ret = numpy.zeros((sum(N), ndim + 2))
    tot = 0
    for state, cnt in enumerate(N):
        for d in range(ndim):
            if lower[d] < upper[d]:
                ret[tot:tot + cnt, d] = (rng._host.uniform(lower[d], upper[d], cnt)
                                         if isinstance(rng, PhiloxRNG)
                                         else rng.uniform(lower[d], upper[d], cnt))
            else:
                ret[tot:tot + cnt, d] = lower[d]
        ret[tot:tot + cnt, ndim] = state
        tot += cnt
    ret[:, ndim + 1] = numpy.arange(ret.shape[0])
    return ret


def sample(t, N, *, lower=None, upper=None, D=None, transmat=None, ndim=3, periodic=False,
           rng=None, boundary=None):
    """Points at each time point: list of (sum(N), ndim + 2) arrays."""
    N = list(N) if isinstance(N, collections.abc.Iterable) else [N]
    lower, upper = _limits(lower, upper, ndim)
    if D is None:
        D = numpy.zeros(len(N))
    elif not isinstance(D, collections.abc.Iterable):
        D = numpy.ones(len(N)) * D
    else:
        D = numpy.asarray(D)
        assert len(D) == len(N)
    if transmat is not None:
        transmat = numpy.asarray(transmat)
        n, m = transmat.shape
        assert n == m
        assert (transmat.diagonal() == 0).all()
        assert n == len(N)
    mode = {None: 1 if periodic else 0, 'none': 0, 'periodic': 1, 'reflect': 2}[boundary]
    if rng is None:
        warnings.warn('A random number generator [rng] is not given.')
        rng = numpy.random.RandomState()
    points = _generate_points(rng, N, lower, upper, ndim)
    n = points.shape[0]
    # moving time points (sampling2.py:136-148)
    moves, tcur = [], t[0]
    for tnext in t[1:]:
        if tnext > tcur:
            moves.append(tnext - tcur)
            tcur = tnext
        else:
            assert tnext == tcur
    T = len(moves)
    nstates = len(N)
    state0 = points[:, ndim].astype(numpy.int32)
    x0 = numpy.ascontiguousarray(points[:, :ndim])
    if isinstance(rng, PhiloxRNG):
        # device draws: per-state D on every axis, transitions after each move
        D_tab = numpy.repeat(numpy.asarray(D, dtype=float)[:, None], ndim, axis=1)
        traj, states = _philox_walk(rng, x0, D_tab, moves, boundary=mode, lower=lower,
                                    upper=upper, state0=state0, transmat=transmat)
    else:
        scale = numpy.empty((T, nstates, ndim))
        xi = numpy.empty((T, n, ndim))
        pacc = tu = None
        if transmat is not None:
            pacc = numpy.empty((T, nstates, nstates))
            tu = numpy.empty((T, n))
        for k, dt in enumerate(moves):
            for s in range(nstates):
                scale[k, s, :] = numpy.sqrt(2 * D[s] * dt)
            xi[k] = rng.standard_normal(n * ndim).reshape(n, ndim)
            if transmat is not None:
                P = 1 - numpy.exp(-transmat * dt)
                assert (P.sum(axis=1) <= 1.0).all()
                P.ravel()[::nstates + 1] = 1.0 - P.sum(axis=1)
                pacc[k] = P.cumsum(axis=1)
                tu[k] = rng.uniform(0, 1, n)
        traj, states = trajectory_predrawn(x0, scale, xi, boundary=mode, lower=lower, upper=upper,
                                           state0=state0, pacc=pacc, trans_u=tu)
    out = [points.copy()]
    k, tcur = 0, t[0]
    cur = points
    for tnext in t[1:]:
        if tnext > tcur:
            k += 1
            tcur = tnext
            cur = points.copy()
            cur[:, :ndim] = traj[k]
            cur[:, ndim] = states[k]
        out.append(cur)
    return out
