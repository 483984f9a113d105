"""Counter-based random numbers for the GPU-native paths.

``PhiloxRNG(seed)`` can be passed wherever scopyon takes ``rng``.  Device
kernels then draw their own numbers from Philox4x32-10 keyed by ``seed`` and
a per-call counter, so nothing is drawn on the host and results are
reproducible for a given seed (but are not the numpy ``RandomState`` stream).
"""
import numpy

__all__ = ['PhiloxRNG']


class PhiloxRNG:

    def __init__(self, seed=0):
        self.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        self.counter = 0
        # host-side draws (rare, e.g. first-sight photobleaching budgets)
        self._host = numpy.random.Generator(numpy.random.Philox(self.seed))

    def next_counter(self):
        c = self.counter
        self.counter += 1
        return c & 0xFFFFFFFF

    def host_exponential(self, scale):
        return float(self._host.exponential(scale))

    def __repr__(self):
        return 'PhiloxRNG(seed={})'.format(self.seed)
