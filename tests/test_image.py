"""Image / Video output formats (src/scopyon/image.py:99-175) against
hand-computed 8-bit values and file round trips."""
import numpy
import pytest

from scopyon_amd.image import Image, Video


def test_as_8bit_default_range():
    # scale 255 / 4: 0 -> 0, 1 -> 63.75 (+0.5) -> 64, 2 -> 127.5 -> 128, 4 -> 255
    img = Image(numpy.array([[0.0, 1.0], [2.0, 4.0]]))
    got = img.as_8bit().as_array()
    assert got.dtype == numpy.uint8
    assert got.tolist() == [[0, 64], [128, 255]]


def test_as_8bit_explicit_range_clips():
    # [1, 3] -> [10, 20]: slope 5; 0 -> 5 (clipped to 10), 2 -> 15, 4 -> 25 (clipped to 20)
    img = Image(numpy.array([[0.0, 1.0], [2.0, 4.0]]))
    got = img.as_8bit(cmin=1.0, cmax=3.0, low=10, high=20).as_array()
    assert got.tolist() == [[10, 10], [15, 20]]


def test_as_8bit_rounds_half_up():
    # range [0, 510]: slope 0.5 exactly; 1 -> 0.5 rounds up, 0.98 -> 0.49 down
    img = Image(numpy.array([[0.0, 1.0, 0.98, 510.0]]))
    assert img.as_8bit().as_array().tolist() == [[0, 1, 0, 255]]


def test_as_8bit_constant_frame_is_low():
    img = Image(numpy.full((3, 2), 7.0))
    assert (img.as_8bit().as_array() == 0).all()
    assert (img.as_8bit(low=3).as_array() == 3).all()


def test_as_8bit_rescales_uint8_too():
    # the reference's uint8 short-cut never fires (dtype `is` numpy.uint8 is False)
    img = Image(numpy.array([[10, 20], [15, 20]], dtype=numpy.uint8))
    assert img.as_8bit().as_array().tolist() == [[0, 255], [128, 255]]


def test_as_8bit_rgb_scales_each_channel():
    r = numpy.array([[0.0, 2.0]])
    g = numpy.array([[5.0, 10.0]])
    img = Image.RGB(r, Image(g))
    assert img.shape == (1, 2, 3)
    got = img.as_8bit().as_array()
    assert got[:, :, 0].tolist() == [[0, 255]]
    assert got[:, :, 1].tolist() == [[0, 255]]
    assert (got[:, :, 2] == 0).all()


def test_rgb_rejects_mismatched_channels():
    with pytest.raises(ValueError):
        Image.RGB(numpy.zeros((2, 2)), numpy.zeros((2, 3)))
    with pytest.raises(ValueError):
        Image.RGB()


def test_image_rejects_bad_shapes():
    with pytest.raises(ValueError):
        Image(numpy.zeros((2, 2, 4)))


def test_save_load_round_trips(tmp_path):
    data = numpy.array([[0.0, 1.5], [2.25, 4.0]])
    img = Image(data)
    for ext in ('.npy', '.csv'):
        p = tmp_path / ('frame' + ext)
        img.save(p)
        assert numpy.array_equal(Image.load(str(p)).as_array(), data)
    p = tmp_path / 'frame.png'
    img.save(str(p))
    assert Image.load(p).as_array().tolist() == [[0, 96], [143, 255]]
    with pytest.raises(TypeError):
        img.save(tmp_path / 'x.png', shapes=[])
    with pytest.raises(TypeError):
        img.save(3)


def test_video_shares_one_scale(tmp_path):
    import PIL.Image
    frames = [Image(numpy.array([[0.0, 1.0]])), numpy.array([[2.0, 4.0]])]
    p = tmp_path / 'movie.tif'
    Video.save(str(p), frames, interval=50)
    with PIL.Image.open(str(p)) as pic:
        assert pic.n_frames == 2
        got = []
        for k in range(2):
            pic.seek(k)
            got.append(numpy.asarray(pic).tolist())
    # global range [0, 4]
    assert got == [[[0, 64]], [[128, 255]]]
