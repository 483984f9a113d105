"""Copy the reference's example output images into compact npz fixtures.

    python tests/golden/make_png_fixture.py   (development container only)

``examples/tirf_000.png`` is what ``examples/tirf.py`` wrote (RandomState(123),
100 emitters, default configuration, 33 ms exposure, bytescale to 8 bit).
Only the pixel arrays are kept.
"""
import os

import numpy
from PIL import Image

SRC = '/root/reference/examples'
HERE = os.path.dirname(os.path.abspath(__file__))

if __name__ == '__main__':
    for name in ('tirf_000', 'twocolor_000'):
        arr = numpy.asarray(Image.open(os.path.join(SRC, name + '.png')))
        numpy.savez_compressed(os.path.join(HERE, name + '_u8.npz'), image=arr)
        print(name, arr.shape, arr.dtype)
