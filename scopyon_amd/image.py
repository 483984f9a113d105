"""``Image``: the camera frame ``form_image`` returns (ADC counts), and
``Video``: a stack of such frames written as one animated file.

Public surface kept from the reference (src/scopyon/image.py:12-238) because
callers use it: ``Image(data)``, ``as_array()``, ``shape`` / ``dtype`` /
``ndim`` / ``size``, ``as_8bit(cmin, cmax, low, high)``, ``save(filename)``,
``Image.load(filename)``, ``Image.RGB(red, green, blue)`` and
``Video.save(filename, images, ...)``.

The 8-bit conversion is the linear ``bytescale`` of image.py:99-123: values
are mapped from [cmin, cmax] (default: the channel's min and max) onto
[low, high] (default 0 and 255), clipped, and rounded half up.  Like the
reference (whose uint8 short-cut at image.py:137 compares a dtype object with
``is`` and never fires), uint8 data is rescaled too.  File formats follow
image.py:152-175: ``.npy`` and ``.csv`` hold the raw values, any other suffix
is an 8-bit picture written by Pillow.  The plotting helpers of the reference
(plotly/matplotlib figures, spot overlays) are not part of this package.
"""
import os
import pathlib

import numpy

__all__ = ["Image", "Video"]


def _path(filename):
    if not isinstance(filename, (str, pathlib.PurePath)):
        raise TypeError('filename must be a str or a path, not {}'.format(type(filename).__name__))
    name = os.fspath(filename)
    return name, os.path.splitext(name)[1].lower()


def _scale_channel(values, cmin, cmax, low, high):
    """One channel onto [low, high] as uint8 (image.py:99-123 arithmetic)."""
    lo_val = values.min() if cmin is None else cmin
    hi_val = values.max() if cmax is None else cmax
    low = 0 if low is None else low
    high = 255 if high is None else high
    span = hi_val - lo_val
    if span == 0.0:
        return numpy.full(values.shape, low, dtype=numpy.uint8)
    k = float(high - low) / span
    mapped = numpy.clip((values - lo_val) * k + low, low, high)
    return (mapped + 0.5).astype(numpy.uint8)


def _to_8bit(data, cmin, cmax, low, high):
    if data.ndim == 2:
        return _scale_channel(data, cmin, cmax, low, high)
    return numpy.stack([_scale_channel(data[:, :, c], cmin, cmax, low, high)
                        for c in range(data.shape[2])], axis=2)


class Image(object):
    """A 2D frame (H x W) or an RGB frame (H x W x 3)."""

    def __init__(self, data):
        data = numpy.asarray(data)
        if not (data.ndim == 2 or (data.ndim == 3 and data.shape[2] == 3)):
            raise ValueError('an Image holds a 2D array or an RGB array (H, W, 3); got shape {}'
                             .format(data.shape))
        self._data = data

    # ---- data access ----
    def as_array(self):
        return self._data

    @property
    def shape(self):
        return self._data.shape

    @property
    def dtype(self):
        return self._data.dtype

    @property
    def ndim(self):
        return self._data.ndim

    @property
    def size(self):
        return self._data.size

    # ---- conversions ----
    def as_8bit(self, cmin=None, cmax=None, low=None, high=None):
        """A new Image of uint8 values: each channel mapped linearly from
        [cmin, cmax] (default: that channel's range) onto [low, high]
        (default [0, 255])."""
        return Image(_to_8bit(self._data, cmin, cmax, low, high))

    @staticmethod
    def RGB(red=None, green=None, blue=None):
        """An RGB Image from up to three channels (2D arrays or Images of one
        shape and dtype); a missing channel is zero."""
        chans = [c.as_array() if isinstance(c, Image) else c for c in (red, green, blue)]
        given = [c for c in chans if c is not None]
        if not given:
            raise ValueError('RGB() needs at least one channel')
        for c in given:
            if not isinstance(c, numpy.ndarray):
                raise TypeError('a channel is a numpy array or an Image')
            if c.shape != given[0].shape or c.dtype != given[0].dtype:
                raise ValueError('RGB channels differ in shape or dtype')
        out = numpy.zeros(given[0].shape[:2] + (3,), dtype=given[0].dtype)
        for k, c in enumerate(chans):
            if c is not None:
                out[:, :, k] = c
        return Image(out)

    # ---- files ----
    def save(self, filename, **kwargs):
        """Write the frame: ``.npy`` (numpy binary) and ``.csv`` (text) keep
        the values; other suffixes (``.png``, ``.tif``, ...) store
        ``as_8bit()`` through Pillow."""
        name, ext = _path(filename)
        if kwargs:
            raise TypeError('Image.save() takes no options here (got {})'.format(sorted(kwargs)))
        if ext == '.npy':
            numpy.save(name, self._data)
        elif ext == '.csv':
            numpy.savetxt(name, self._data)
        else:
            import PIL.Image
            PIL.Image.fromarray(self.as_8bit().as_array()).save(name)

    @staticmethod
    def load(filename):
        """Read a frame written by ``save`` (or any picture Pillow reads)."""
        name, ext = _path(filename)
        if ext == '.npy':
            return Image(numpy.load(name))
        if ext == '.csv':
            return Image(numpy.loadtxt(name))
        import PIL.Image
        with PIL.Image.open(name) as pic:
            return Image(numpy.asarray(pic))


class Video(object):
    """Frames written as one animated picture (GIF, multi-page TIFF, ...)."""

    @staticmethod
    def save(filename, imgs, interval=100, dpi=None, cmin=None, cmax=None, low=None, high=None):
        """Write `imgs` (Images or arrays) with one 8-bit scale shared by all
        frames (default: the global min and max), `interval` ms per frame.
        `dpi` is accepted for call compatibility; frames keep their pixel
        size."""
        del dpi
        name, _ = _path(filename)
        arrays = [i.as_array() if isinstance(i, Image) else numpy.asarray(i) for i in imgs]
        if not arrays:
            raise ValueError('Video.save() needs at least one frame')
        lo = min(a.min() for a in arrays) if cmin is None else cmin
        hi = max(a.max() for a in arrays) if cmax is None else cmax
        import PIL.Image
        pics = [PIL.Image.fromarray(_to_8bit(a, lo, hi, low, high)) for a in arrays]
        pics[0].save(name, save_all=True, append_images=pics[1:], duration=int(interval), loop=0)
