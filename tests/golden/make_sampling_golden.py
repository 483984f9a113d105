"""Write golden vectors for the Brownian samplers from the reference itself.

Run in the development container only (it reads the scopyon sources under
/root/reference, which never travel with the repository):

    python tests/golden/make_sampling_golden.py

``sampling.py`` and ``sampling2.py`` import only numpy/logging/warnings, so
they are loaded by file path without the package ``__init__`` (which needs
pint and hmmlearn).  The output ``sampling_golden.npz`` holds inputs and the
reference's outputs for small seeded cases; nothing else is kept.
"""
import importlib.util
import os
import warnings

import numpy

SRC = '/root/reference/src/scopyon'
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'sampling_golden.npz')


def _load(name):
    spec = importlib.util.spec_from_file_location('ref_' + name, os.path.join(SRC, name + '.py'))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def main():
    s1 = _load('sampling')
    s2 = _load('sampling2')
    out = {}
    # sampling.sample_inputs: 2D, per-axis scalar D, 30 points, 6 time points
    t = numpy.arange(0, 6) * 0.033
    res = s1.sample_inputs(t, N=30, lower=-15e-6, upper=15e-6, ndim=2, D=0.1e-12,
                           rng=numpy.random.RandomState(123))
    out['inputs2d_t'] = numpy.array([a for a, _ in res])
    out['inputs2d_x'] = numpy.stack([b for _, b in res])
    # sampling.sample_inputs: 3D, per-axis D vector, unsorted t, start id 7
    t3 = [0.02, 0.0, 0.01, 0.01, 0.05]
    res = s1.sample_inputs(t3, N=12, lower=[0, 0, -1e-6], upper=[5e-6, 4e-6, 1e-6], ndim=3,
                           D=[1e-12, 2e-12, 0.5e-12], start=7, rng=numpy.random.RandomState(7))
    out['inputs3d_t'] = numpy.array([a for a, _ in res])
    out['inputs3d_x'] = numpy.stack([b for _, b in res])
    # sampling2.sample: 3 states, per-state D, periodic box, no transitions
    ts = numpy.array([0.0, 0.1, 0.2, 0.2, 0.35])
    res = s2.sample(ts, N=[5, 3, 2], lower=[0.0, -1.0], upper=[1.0, 1.0], D=[4e-3, 1e-3, 0.0],
                    ndim=2, periodic=True, rng=numpy.random.RandomState(5))
    out['sample_periodic_x'] = numpy.stack(res)
    # sampling2.sample: with a transition-rate matrix (side='leff' path)
    with warnings.catch_warnings():
        warnings.simplefilter('ignore')
        res = s2.sample([0.0, 0.1, 0.2, 0.3], N=[5, 3, 2], D=[4e-3, 1e-3, 0.0],
                        transmat=[[0.0, 1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]],
                        ndim=2, rng=numpy.random.RandomState(11))
    out['sample_trans_x'] = numpy.stack(res)
    # sampling2.sample: 3D, one state, reversed bounds, no boundary
    res = s2.sample([0.0, 0.5, 1.0], N=8, lower=[1.0, 1.0, 1.0], upper=[0.0, 0.0, 0.5],
                    D=0.01, ndim=3, rng=numpy.random.RandomState(3))
    out['sample_3d_x'] = numpy.stack(res)
    numpy.savez_compressed(OUT, **out)
    print('wrote', OUT, sorted(out))


if __name__ == '__main__':
    main()
