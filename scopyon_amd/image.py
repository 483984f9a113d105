"""Image container returned by ``form_image`` (src/scopyon/image.py:12-238).

Keeps the reference's surface for what the hot path returns: ``as_array``,
``as_8bit`` (bytescale), ``save`` (.npy / .csv / image formats through PIL),
``load``, ``RGB`` and the shape/dtype properties.  ``Video.save`` writes an
animation through matplotlib when it is installed.
"""
import os
import pathlib

import numpy

__all__ = ["Image", "Video"]


class Image(object):

    @staticmethod
    def load(file):
        assert isinstance(file, (str, pathlib.PurePath))
        filename = str(file)
        ext = os.path.splitext(filename)[1].lower()
        if ext == '.npy':
            return Image(numpy.load(filename))
        if ext == '.csv':
            return Image(numpy.loadtxt(filename))
        import PIL.Image
        return Image(numpy.asarray(PIL.Image.open(filename)))

    @staticmethod
    def RGB(red=None, green=None, blue=None):
        shape, dtype = None, None
        for ch in (red, green, blue):
            if ch is None:
                continue
            if shape is None:
                shape, dtype = ch.shape, ch.dtype
            else:
                assert shape == ch.shape and dtype == ch.dtype
        assert shape is not None
        rgb = numpy.zeros((shape[0], shape[1], 3), dtype=dtype)
        for i, ch in enumerate((red, green, blue)):
            if isinstance(ch, Image):
                rgb[:, :, i] = ch.as_array()
            elif isinstance(ch, numpy.ndarray):
                rgb[:, :, i] = ch
            else:
                assert ch is None
        return Image(rgb)

    def __init__(self, data):
        assert data.ndim == 2 or (data.ndim == 3 and data.shape[2] == 3)
        self.__data = data

    def as_array(self):
        return self.__data

    @property
    def dtype(self):
        return self.__data.dtype

    @property
    def ndim(self):
        return self.__data.ndim

    @property
    def size(self):
        return self.__data.size

    @property
    def shape(self):
        return self.__data.shape

    @staticmethod
    def _bytescale(data, cmin=None, cmax=None, low=None, high=None):
        """image.py:99-123."""
        cmin = data.min() if cmin is None else cmin
        cmax = data.max() if cmax is None else cmax
        low = 0 if low is None else low
        high = 255 if high is None else high
        cscale = cmax - cmin
        if cscale == 0.0:
            return numpy.ones(data.shape, dtype=numpy.uint8) * low
        scale = float(high - low) / cscale
        byte = (data - cmin) * scale + low
        return (byte.clip(low, high) + 0.5).astype(numpy.uint8)

    def as_8bit(self, cmin=None, cmax=None, low=None, high=None):
        if self.dtype == numpy.uint8:
            data = self.__data.copy()
        elif self.ndim == 2:
            data = self._bytescale(self.__data, cmin, cmax, low, high)
        else:
            data = numpy.zeros(self.__data.shape, dtype=numpy.uint8)
            for i in range(3):
                data[:, :, i] = self._bytescale(self.__data[:, :, i], cmin, cmax, low, high)
        return Image(data)

    def save(self, filename, **kwargs):
        assert isinstance(filename, (str, pathlib.PurePath))
        filename = str(filename)
        ext = os.path.splitext(filename)[1].lower()
        if ext == '.npy':
            assert not kwargs
            numpy.save(filename, self.__data)
        elif ext == '.csv':
            assert not kwargs
            numpy.savetxt(filename, self.__data)
        else:
            self.savefig(filename, self.as_8bit().as_array(), **kwargs)

    @staticmethod
    def savefig(filename, img, shapes=None):
        import PIL.Image
        from PIL import ImageDraw
        pic = PIL.Image.fromarray(img)
        if shapes is not None:
            pic = pic.convert('RGB')
            d = ImageDraw.Draw(pic)
            for s in shapes:
                x, y = s['y'], s['x']
                d.rectangle([(x - s['sigma'], y - s['sigma']), (x + s['sigma'], y + s['sigma'])],
                            outline=s['color'], width=1)
        pic.save(filename)

    def show(self, **kwargs):
        import matplotlib.pyplot as plt
        plt.imshow(self.as_8bit().as_array(), cmap='gray', interpolation='none', **kwargs)
        plt.show()


class Video:

    @staticmethod
    def save(filename, imgs, interval=100, dpi=None, cmin=None, cmax=None, low=None, high=None):
        import matplotlib.pyplot as plt
        from matplotlib.animation import FuncAnimation
        if cmin is None:
            cmin = min(img.as_array().min() for img in imgs)
        if cmax is None:
            cmax = max(img.as_array().max() for img in imgs)
        plt.ioff()
        fig, ax = plt.subplots(1, figsize=(1, 1))
        fig.subplots_adjust(0, 0, 1, 1)
        ax.axis('off')
        art = ax.imshow(imgs[0].as_8bit(cmin, cmax, low, high).as_array(), cmap='gray')

        def animate(k):
            art.set_array(imgs[k].as_8bit(cmin, cmax, low, high).as_array())
            return art
        anim = FuncAnimation(fig, animate, numpy.arange(len(imgs)), interval=interval)
        anim.save(filename, dpi=dpi or max(imgs[0].shape))
        plt.close(fig)
