"""Public image-formation API (src/scopyon/base.py:20-294), backed by the GPU
pipeline in ``_epifm``.  Signatures, input formats and errors are the
reference's."""
import collections.abc
import numbers
import warnings

import numpy

from . import _epifm, config as _config, image as _image

from logging import getLogger
_log = getLogger(__name__)

__all__ = ["EnvironSettings", "EPIFMSimulator", "form_image", "generate_images",
           "create_simulator"]


class EnvironSettings:

    def __init__(self, config):
        self.initialize(config)

    def initialize(self, config):
        self.processes = config.processes


class EPIFMSimulator(object):

    def __init__(self, config=None, method=None, rng=None):
        if config is None:
            config = _config.Configuration()
        elif isinstance(config, str):
            config = _config.Configuration(filename=config)
        elif not isinstance(config, _config.Configuration):
            raise TypeError("Configuration or str must be given [{}].".format(type(config)))
        if rng is None:
            warnings.warn('A random number generator is not given.')
            rng = numpy.random.RandomState()
        self.__config = config
        self.__method = method or config.default.lower()
        self.__rng = rng

    def base(self):
        return _epifm._EPIFMSimulator(
            configs=_epifm.EPIFMConfigs(self.__config[self.__method], rng=self.__rng),
            environ=EnvironSettings(self.__config.environ))

    def __format_data(self, inputs):
        """base.py:61-110 -> (n, 5) rows (depth, u, v, m_id, p_state)."""
        assert isinstance(inputs, numpy.ndarray)
        if inputs.ndim != 2:
            raise ValueError("The given 'inputs' has wrong dimension.")
        pre = self.__config.preprocessing
        n, ncol = inputs.shape
        if ncol == 2:
            data = numpy.hstack((numpy.zeros((n, 1)), inputs * pre.scale))
            data = numpy.hstack((data, numpy.zeros((n, 2))))
            data[:, 3] = numpy.arange(n)
            data[:, 4] = 1.0
        elif ncol in (3, 5):
            origin = numpy.array(pre.origin)
            xyz = inputs[:, :3] * pre.scale - origin
            unit_z = numpy.cross(pre.unit_x, pre.unit_y)
            cols = [numpy.dot(xyz, unit_z).reshape((-1, 1)),
                    numpy.dot(xyz, pre.unit_x).reshape((-1, 1)),
                    numpy.dot(xyz, pre.unit_y).reshape((-1, 1))]
            if ncol == 5:
                data = numpy.hstack(cols + [inputs[:, 3:]])
            else:
                data = numpy.hstack(cols + [numpy.zeros((n, 2))])
                data[:, 3] = numpy.arange(n)
                data[:, 4] = 1.0
        elif ncol == 4:
            data = numpy.hstack((numpy.zeros((n, 1)), inputs[:, :2] * pre.scale, inputs[:, 2:]))
        else:
            raise ValueError("The given 'inputs' has wrong shape.")
        return data

    def __format_inputs(self, inputs):
        """base.py:112-127."""
        if isinstance(inputs, numpy.ndarray):
            return ((0.0, self.__format_data(inputs)), )
        if isinstance(inputs, collections.abc.Iterable):
            data = []
            for elem in inputs:
                if not (isinstance(elem, (tuple, list)) and len(elem) == 2
                        and isinstance(elem[0], numbers.Real) and isinstance(elem[1], numpy.ndarray)):
                    raise ValueError("The given 'inputs' has wrong type.")
                data.append((elem[0], self.__format_data(elem[1])))
            return data
        raise TypeError("Invalid argument was given [{}]. A ndarray is expected.".format(type(inputs)))

    def form_image(self, inputs, start_time=0.0, exposure_time=None, full_output=False):
        """One frame; returns Image (and the info dict when full_output)."""
        data = self.__format_inputs(inputs)
        base = self.base()
        camera, infodict = base.output_frame(data, start_time=start_time,
                                             exposure_time=exposure_time, rng=self.__rng,
                                             full_output=full_output)
        img = _image.Image(camera[:, :, 1])
        if full_output:
            infodict.update(dict(expectation=camera[:, :, 0]))
            return img, infodict
        return img

    def generate_images(self, inputs, num_frames, start_time=0.0, exposure_time=None,
                        full_output=False):
        """A generator of frames; photobleaching state persists across them."""
        data = self.__format_inputs(inputs)
        base = self.base()
        for camera, infodict in base.generate_frames(data, num_frames, start_time=start_time,
                                                     exposure_time=exposure_time, rng=self.__rng,
                                                     full_output=full_output):
            img = _image.Image(camera[:, :, 1])
            if full_output:
                infodict.update(dict(expectation=camera[:, :, 0]))
                yield img, infodict
            else:
                yield img


def create_simulator(config=None, method=None, rng=None):
    """base.py:203-225."""
    if method is None:
        method = config.get('method', 'default').lower()
    simulator_type = config[method].type.lower()
    if simulator_type == 'epifm':
        return EPIFMSimulator(config=config, method=method, rng=rng)
    raise ValueError(f"An unknown type [{simulator_type}] was given.")


def form_image(inputs, start_time=0.0, exposure_time=None, *, method=None, config=None, rng=None,
               full_output=False):
    """base.py:227-258."""
    sim = create_simulator(config, method=method, rng=rng)
    return sim.form_image(inputs, start_time, exposure_time, full_output=full_output)


def generate_images(inputs, num_frames, start_time=0.0, exposure_time=None, *, method=None,
                    config=None, rng=None, full_output=False):
    """base.py:260-294."""
    sim = create_simulator(config, method=method, rng=rng)
    return sim.generate_images(inputs, num_frames, start_time, exposure_time,
                               full_output=full_output)
