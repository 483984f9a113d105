"""Pins for the oracle's pixel integration (overlay_signal_)."""
import math

import numpy
import pytest

from oracle import epifm as O

RES = 1e-9


def _brute(shape, table, xi, yi, pl, w):
    """Cell by cell: table cell (a, b) sits at object coordinate
    (o_u + a res, o_v + b res) and belongs to pixel (floor(x / pl), floor(y / pl))
    of the frame [0, nw pl) x [0, nh pl); sum the cells of each pixel."""
    nw, nh = shape
    ext = RES * (table.shape[0] - 1)
    ou = (nw * pl * 0.5 + xi) - ext * 0.5
    ov = (nh * pl * 0.5 + yi) - ext * 0.5
    iu = numpy.floor((ou + numpy.arange(table.shape[0]) * RES) / pl).astype(int)
    iv = numpy.floor((ov + numpy.arange(table.shape[1]) * RES) / pl).astype(int)
    acc = numpy.zeros((nw + 2, nh + 2))
    ku = numpy.clip(iu, -1, nw) + 1
    kv = numpy.clip(iv, -1, nh) + 1
    numpy.add.at(acc, (ku[:, None].repeat(table.shape[1], 1), kv[None, :].repeat(table.shape[0], 0)),
                 table)
    return acc[1:-1, 1:-1] * RES * RES * w


def _small_table(n=60):
    r = numpy.arange(n) * RES
    prof = numpy.exp(-(r / (15 * RES)) ** 2) + 0.01
    return O.radial_to_cartesian(r, prof)


@pytest.mark.parametrize('seed', range(6))
def test_overlay_matches_brute_force(seed):
    rng = numpy.random.RandomState(seed)
    table = _small_table()
    pl = [7e-9, 10e-9, 13.3e-9][seed % 3]
    shape = (9, 12)
    xi, yi = rng.uniform(-60e-9, 60e-9, 2)
    exp = numpy.zeros(shape)
    O.overlay_signal_(exp, table, (0.0, xi, yi), pl, 2.5)
    ref = _brute(shape, table, xi, yi, pl, 2.5)
    numpy.testing.assert_allclose(exp, ref, rtol=1e-12, atol=1e-300)


def test_overlay_conserves_photons_inside_frame():
    table = _small_table()
    exp = numpy.zeros((40, 40))
    O.overlay_signal_(exp, table, (0.0, 3e-9, -4e-9), 10e-9, 1.0)
    assert exp.sum() == pytest.approx(table.sum() * RES * RES, rel=1e-12)


def test_overlay_translation_by_whole_pixels():
    table = _small_table()
    pl = 10e-9
    a = numpy.zeros((30, 30))
    b = numpy.zeros((30, 30))
    O.overlay_signal_(a, table, (0.0, 1.3e-9, 2.1e-9), pl, 1.0)
    O.overlay_signal_(b, table, (0.0, 1.3e-9 + 3 * pl, 2.1e-9 - 2 * pl), pl, 1.0)
    numpy.testing.assert_allclose(b[3:, :-2], a[:-3, 2:], rtol=1e-9, atol=1e-30)


def test_overlay_planted_mistake_is_caught(monkeypatch):
    """Plant the (nw - 1) * pl frame-width convention (the one the reference
    explicitly rejects, _epifm.py:230) in the oracle: the brute force catches it."""
    table = _small_table()
    shape, pl = (9, 12), 10e-9
    ref = _brute(shape, table, 4.37e-9, -3.21e-9, pl, 1.0)
    good = numpy.zeros(shape)
    O.overlay_signal_(good, table, (0.0, 4.37e-9, -3.21e-9), pl, 1.0)
    numpy.testing.assert_allclose(good, ref, rtol=1e-12, atol=0)
    orig = O._boundaries

    def shifted(n_pix, pixel_length, offset, signal_len, extent):
        return orig(n_pix - 1, pixel_length, offset, signal_len, extent)
    monkeypatch.setattr(O, '_boundaries', shifted)
    bad = numpy.zeros(shape)
    O.overlay_signal_(bad, table, (0.0, 4.37e-9, -3.21e-9), pl, 1.0)
    assert not numpy.allclose(bad, ref, rtol=1e-6, atol=0)
