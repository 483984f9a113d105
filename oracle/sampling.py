"""float64 restatement of scopyon's Brownian samplers.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  Random numbers are
consumed from a ``numpy.random.RandomState`` in exactly the reference's order
(group -> dim -> particle for the initial uniforms, particle -> dim for the
Brownian normals), so the same seed gives the same trajectories.
"""
import collections.abc
import numbers

import numpy


def _limits(lower, upper, ndim):
    lower = (numpy.ones(ndim) * lower if isinstance(lower, numbers.Number)
             else numpy.array(lower) if lower is not None else numpy.zeros(ndim))
    upper = (numpy.ones(ndim) * upper if isinstance(upper, numbers.Number)
             else numpy.array(upper) if upper is not None else numpy.ones(ndim))
    if len(lower) < ndim or len(upper) < ndim:
        raise ValueError("The wrong size of limits was given")
    for d in range(ndim):
        if lower[d] > upper[d]:
            lower[d], upper[d] = upper[d], lower[d]
    return lower, upper


# ---------------------------------------------------------------- sampling2.py

def generate_points(rng, N, lower, upper, ndim):
    """sampling2.py:15-31: columns [coords..., state, molecule id]."""
    ret = numpy.zeros((sum(N), ndim + 2))
    tot = 0
    for state, cnt in enumerate(N):
        for d in range(ndim):
            if lower[d] < upper[d]:
                ret[tot:tot + cnt, d] = rng.uniform(lower[d], upper[d], cnt)
            else:
                ret[tot:tot + cnt, d] = lower[d]
        ret[tot:tot + cnt, ndim] = state
        tot += cnt
    ret[:, ndim + 1] = numpy.arange(ret.shape[0])
    return ret


def move_points(rng, points, D, lower, upper, dt, ndim, periodic):
    """sampling2.py:33-50: per-state D, scalar normal per (particle, dim)."""
    ret = points.copy()
    for i in range(len(ret)):
        scale = numpy.sqrt(2 * D[int(ret[i, ndim])] * dt)
        for d in range(ndim):
            ret[i, d] += rng.normal(0.0, scale)
    if periodic:
        for d in range(ndim):
            if upper[d] > lower[d]:
                ret[:, d] = (ret[:, d] - lower[d]) % (upper[d] - lower[d]) + lower[d]
            else:
                ret[:, d] = lower[d]
    return ret


def transition_states(rng, points, transmat, dt, ndim):
    """sampling2.py:52-68 (``side='leff'`` behaves as 'left' on numpy 2.x)."""
    n = transmat.shape[0]
    P = 1 - numpy.exp(-transmat * dt)
    assert (P.sum(axis=1) <= 1.0).all()
    P.ravel()[::n + 1] = 1.0 - P.sum(axis=1)
    Pacc = P.cumsum(axis=1)
    ret = points.copy()
    for i in range(ret.shape[0]):
        state = int(ret[i, ndim])
        u = rng.uniform(0, 1)
        ret[i, ndim] = numpy.searchsorted(Pacc[state], u, side='left')
    return ret


def sample(t, N, *, lower=None, upper=None, D=None, transmat=None, ndim=3,
           periodic=False, rng=None):
    """sampling2.py:70-149."""
    N = list(N) if isinstance(N, collections.abc.Iterable) else [N]
    lower, upper = _limits(lower, upper, ndim)
    if D is None:
        D = numpy.zeros(len(N))
    elif not isinstance(D, collections.abc.Iterable):
        D = numpy.ones(len(N)) * D
    else:
        D = numpy.asarray(D)
    if transmat is not None:
        transmat = numpy.asarray(transmat)
    points = generate_points(rng, N, lower, upper, ndim)
    tcur = t[0]
    out = [points.copy()]
    for tnext in t[1:]:
        if tnext > tcur:
            dt = tnext - tcur
            points = move_points(rng, points, D, lower, upper, dt, ndim, periodic)
            if transmat is not None:
                points = transition_states(rng, points, transmat, dt, ndim)
            tcur = tnext
        else:
            assert tnext == tcur
        out.append(points)
    return out


# ----------------------------------------------------------------- sampling.py

def sample_points(rng, N=None, conc=None, lower=None, upper=None, start=0, ndim=3):
    """sampling.py:14-86: columns [coords..., molecule id, p_state=1]."""
    if N is None and conc is None:
        raise ValueError('Either one of N or conc must be given.')
    lower, upper = _limits(lower, upper, ndim)
    if N is None:
        lengths = upper - lower
        size = numpy.prod(lengths[lengths != 0])
        concs = conc if isinstance(conc, collections.abc.Iterable) else [conc]
        counts = [rng.poisson(size * c) for c in concs]
    elif not isinstance(N, collections.abc.Iterable):
        counts = [N]
    else:
        counts = N
    N = sum(counts)
    if N <= 0:
        return numpy.array([]), start
    maxdim = len(lower)
    ret = numpy.zeros((N, maxdim + 2))
    for d in range(maxdim):
        if lower[d] < upper[d]:
            ret[:, d] = rng.uniform(lower[d], upper[d], N)
        else:
            ret[:, d] = lower[d]
    ret[:, maxdim] = numpy.arange(start, start + N)
    ret[:, maxdim + 1] = 1.0
    return ret, start + N


def move_points_axes(rng, points, D, dt, ndim=3):
    """sampling.py:88-119: per-axis D."""
    if D is None:
        D = numpy.zeros(ndim)
    elif not isinstance(D, collections.abc.Iterable):
        D = numpy.ones(ndim) * D
    else:
        D = numpy.asarray(D)
    ret = points.copy()
    for i in range(len(ret)):
        for d in range(ndim):
            ret[i, d] += rng.normal(0.0, numpy.sqrt(2 * D[d] * dt))
    return ret


def sample_inputs(t, *, N=None, conc=None, lower=None, upper=None, D=None, start=0,
                  ndim=3, rng=None):
    """sampling.py:121-162 -> [(t_k, points_k)]."""
    t = sorted(t)
    points, _ = sample_points(rng, N=N, conc=conc, lower=lower, upper=upper, start=start,
                              ndim=ndim)
    tcur = t[0]
    out = []
    for tnext in t:
        if tnext > tcur:
            points = move_points_axes(rng, points, D=D, dt=(tnext - tcur), ndim=ndim)
            tcur = tnext
        else:
            points = points.copy()
        out.append((tcur, points))
    return out
