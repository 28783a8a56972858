"""The localization-bitmap oracle (oracle/lzr_render_oracle.c), pinned by what the
mathematics fixes: the erf table against math.erf, the telescoping sum of one emitter's
pixels, translation by one output pixel, mirror symmetry, exact additivity in position
order, the three deviation calculators and the two weighings against each other, and the
reference's error cases (source/PALMImages.cpp:68-177).  The reference renderer needs
GSL's erf (libgsl is not installed), so it cannot be built here."""
import math

import numpy as np
import pytest

from oracle import pyoracle as po

SAME, INTEGRAL = 0, 1
CONSTANT, FIT_UNC, MASK = 0, 1, 2


def test_erf_table_matches_erf_within_interpolation_error():
    t = po.erf_table()
    assert len(t) == 80001 + 1
    rng = np.random.default_rng(1)
    for x in rng.uniform(-4.0, 4.0, 2000):
        # linear interpolation error <= h^2/8 max|erf''| = 1e-8/8 * 0.97
        assert abs(po.erf_lookup(t, x) - math.erf(x)) < 1.3e-9
    for i in (0, 1, 12345, 40000, 79999):
        assert po.erf_lookup(t, -4.0 + i * 1e-4) == pytest.approx(math.erf(-4.0 + i * 1e-4), abs=1e-15)
    assert po.erf_lookup(t, -4.0000001) == -1.0 and po.erf_lookup(t, 4.0000001) == 1.0
    assert po.erf_lookup(t, 4.0) == t[80000]


def _one(x, y, dev, x_size=9, y_size=9, scale=33.0001 / 9, integral=1.0, weighing=SAME):
    return po.render(np.array([[x, y, integral, 0.0, 0.0]]), x_size, y_size, scale, CONSTANT, [dev], weighing)


def test_single_emitter_telescopes_to_the_box_integral():
    """Summed over the pixel box, 0.25 (E(maxX) - E(minX)) (E(maxY) - E(minY)) telescopes
    to the box edges (consecutive pixels share an edge: maxX_i = minX_{i+1})."""
    t = po.erf_table()
    x_size, scale = 9, 33.0001 / 9
    w, h = po.render_size(x_size, x_size, scale)
    sf = (w - 1) / (x_size - 1)
    assert sf == 4.0
    sps, half = 1 / sf, 0.5 / sf
    for x0, y0, dev in ((3.3, 4.1, 0.3), (4.0, 4.0, 0.25), (2.7, 5.55, 0.4)):
        img = _one(x0, y0, dev)
        s = math.sqrt(2 * dev * dev)
        r0, r1 = math.floor(x0 * sf - 4 * dev * sf), math.ceil(x0 * sf + 4 * dev * sf)
        c0, c1 = math.floor(y0 * sf - 4 * dev * sf), math.ceil(y0 * sf + 4 * dev * sf)
        ex = po.erf_lookup(t, (r1 * sps - half + sps - x0) / s) - po.erf_lookup(t, (r0 * sps - half - x0) / s)
        ey = po.erf_lookup(t, (c1 * sps - half + sps - y0) / s) - po.erf_lookup(t, (c0 * sps - half - y0) / s)
        assert img.sum() == pytest.approx(0.25 * ex * ey, rel=1e-12)
        # nothing outside the box, and the box holds ~all of a +-4 sigma Gaussian
        mask = np.zeros_like(img, bool)
        mask[c0:c1 + 1, r0:r1 + 1] = True
        assert np.all(img[~mask] == 0.0)
        assert img.sum() == pytest.approx(math.erf(4 / math.sqrt(2)) ** 2, abs=2e-4)


def test_translation_by_one_output_pixel():
    """sf = 4: moving the emitter by 1/4 input px moves the image by one pixel exactly."""
    a = _one(3.25, 4.5, 0.3)
    b = _one(3.5, 4.5, 0.3)
    c = _one(3.25, 4.75, 0.3)
    assert np.array_equal(b[:, 1:], a[:, :-1])
    assert np.array_equal(c[1:, :], a[:-1, :])


def test_mirror_symmetry_about_the_image_centre():
    img = _one(4.0, 4.0, 0.35)  # pixel i covers [i/4 - 1/8, i/4 + 1/8): centre at 4.0 <-> i = 16
    assert img.shape == (33, 33)
    np.testing.assert_allclose(img, img[:, ::-1], rtol=0, atol=1e-15)
    np.testing.assert_allclose(img, img[::-1, :], rtol=0, atol=1e-15)
    np.testing.assert_allclose(img, img.T, rtol=0, atol=1e-15)


def test_additive_in_position_order():
    rng = np.random.default_rng(3)
    pos = np.c_[rng.uniform(1, 30, 5), rng.uniform(1, 20, 5), rng.uniform(100, 900, 5), rng.uniform(0.01, 0.1, 5),
                rng.uniform(50, 150, 5)]
    full = po.render(pos, 32, 24, 3.0, CONSTANT, [0.4], INTEGRAL)
    acc = np.zeros_like(full)
    for k in range(5):
        acc = acc + po.render(pos[k:k + 1], 32, 24, 3.0, CONSTANT, [0.4], INTEGRAL)
    assert np.array_equal(full, acc)  # the reference's += in position order, bit for bit


def test_weighing_and_deviation_calculators():
    pos = np.array([[10.2, 7.7, 640.0, 0.03, 120.0]])
    same = po.render(pos, 20, 16, 4.0, CONSTANT, [0.3], SAME)
    integ = po.render(pos, 20, 16, 4.0, CONSTANT, [0.3], INTEGRAL)
    assert np.array_equal(integ, same * 640.0)
    # fit uncertainty: dev = xDev * scale below the limit, nothing at or above it
    fu = po.render(pos, 20, 16, 4.0, FIT_UNC, [10.0, 0.05], SAME)
    assert np.array_equal(fu, po.render(pos, 20, 16, 4.0, CONSTANT, [0.03 * 10.0], SAME))
    assert not po.render(pos, 20, 16, 4.0, FIT_UNC, [10.0, 0.03], SAME).any()
    # Gaussian mask: dev = sqrt((w^2 + 1/12) / I'), I' = floor(I/m + 0.5) (Mortensen)
    m = 2.5
    Ip = math.floor(640.0 / m + 0.5)
    dev = math.sqrt((1.3 * 1.3 + 1.0 / 12.0) / Ip)
    gm = po.render(pos, 20, 16, 4.0, MASK, [1.3, 100.0, m], SAME)
    assert np.array_equal(gm, po.render(pos, 20, 16, 4.0, CONSTANT, [dev], SAME))
    # zero photons after rounding: infinite deviation, no pixel (the x86-64 conversion)
    assert not po.render(np.array([[10.2, 7.7, 1.0, 0.03, 120.0]]), 20, 16, 4.0, MASK, [1.3, 0.0, 5.0], SAME).any()


def test_skips_and_errors():
    empty = po.render(np.array([[10.0, 8.0, 0.0, 0.1, 1.0]]), 20, 16, 4.0, CONSTANT, [0.3], INTEGRAL)
    assert not empty.any()  # integral 0 skipped under integral weighing
    assert not po.render(np.array([[np.nan, 8.0, 5.0, 0.1, 1.0]]), 20, 16, 4.0, CONSTANT, [0.3], SAME).any()
    # emitters partly or wholly outside the image are clipped, never an error
    img = po.render(np.array([[-0.2, 0.1, 1.0, 0.0, 0.0], [25.0, 3.0, 1.0, 0.0, 0.0]]), 20, 16, 4.0, CONSTANT,
                    [0.3], SAME)
    assert 0.0 < img.sum() < 1.0
    with pytest.raises(ValueError):
        po.render(np.zeros((1, 5)), 20, 16, 0.0, CONSTANT, [0.3], SAME)  # outputScaleFactor <= 0
    with pytest.raises(ValueError):
        po.render(np.zeros((1, 5)), 20, 16, 2.0, CONSTANT, [0.3], 2)  # unknown weighing
    with pytest.raises(ValueError):  # Gaussian mask needs an integral and a background
        po.render(np.array([[3.0, 3.0, -1.0, 0.1, 5.0]]), 20, 16, 2.0, MASK, [1.3, 0.0, 1.0], SAME)
