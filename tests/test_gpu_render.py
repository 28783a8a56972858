"""Localization bitmap on the GPU (lzr_render_device) against the oracle
(oracle/lzr_render_oracle.c, PALMBitmapImageCalculator::CalculateImage,
source/PALMImages.cpp:68-157): the same erf table and the same position-ordered FP64
sums, so the images must be identical bit for bit."""
import sys

import numpy as np
import pytest

from localizer_amd import abi, device, native
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _positions(rng, n, x_size, y_size, edge=True):
    pos = np.zeros(n, abi.POSITION_DTYPE)
    lo = -1.0 if edge else 1.0
    pos["x"] = rng.uniform(lo, x_size - 1 - lo, n)
    pos["y"] = rng.uniform(lo, y_size - 1 - lo, n)
    pos["integral"] = rng.uniform(50, 5000, n)
    pos["x_deviation"] = rng.uniform(0.005, 0.12, n)
    pos["background"] = rng.uniform(80, 200, n)
    return pos


def _as_rows(pos):
    return np.c_[pos["x"], pos["y"], pos["integral"], pos["x_deviation"], pos["background"]]


def _gpu(r, pos, x_size, y_size, scale, method, params, weighing):
    w, h = native.render_image_size(x_size, y_size, scale)
    d_pos = device.upload(pos) if len(pos) else device.empty(abi.POSITION_DTYPE, 0)
    d_img = device.empty(np.float64, w * h)
    r.render_device(d_pos.ptr, len(pos), x_size, y_size, scale, method, params, weighing, d_img.ptr,
                    device.stream_handle())
    return d_img.numpy().reshape(h, w)


CASES = [
    # (name, n, x_size, y_size, scale, method, params, weighing)
    ("constant, same weight", 400, 40, 30, 5.0, abi.DEVIATION_CONSTANT, [0.25], abi.WEIGHING_SAME),
    ("constant, integral weight", 400, 40, 30, 7.3, abi.DEVIATION_CONSTANT, [0.18], abi.WEIGHING_INTEGRAL),
    ("fit uncertainty", 600, 33, 47, 6.0, abi.DEVIATION_FIT_UNCERTAINTY, [3.0, 0.09], abi.WEIGHING_INTEGRAL),
    ("Gaussian mask", 500, 50, 20, 4.0, abi.DEVIATION_GAUSSIAN_MASK, [1.3, 100.0, 2.0], abi.WEIGHING_SAME),
    ("odd scale, large deviation", 200, 17, 23, 2.9, abi.DEVIATION_CONSTANT, [1.7], abi.WEIGHING_SAME),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_gpu_render_matches_oracle(case):
    name, n, xs, ys, scale, method, params, weighing = CASES[case]
    _say(f"case: render {name}: {n} positions, {xs}x{ys} px x {scale}")
    rng = np.random.default_rng(100 + case)
    pos = _positions(rng, n, xs, ys)
    pos["integral"][::37] = 0.0  # skipped under integral weighing
    pos["x"][::53] = np.nan      # no box
    with native.Renderer() as r:
        got = _gpu(r, pos, xs, ys, scale, method, params, weighing)
    want = po.render(_as_rows(pos), xs, ys, scale, method, params, weighing)
    assert got.shape == want.shape
    assert np.array_equal(got, want)
    assert want.any()


def test_gpu_render_dense_tiles_and_empty_input():
    """Thousands of positions on one spot: a tile's list is staged through LDS in many
    chunks of 256 and must keep position order."""
    _say("case: render 3000 positions on a 6x6 px area (chunked tile lists), and an empty input")
    rng = np.random.default_rng(7)
    pos = _positions(rng, 3000, 6, 6, edge=False)
    pos["x"] = rng.normal(3.0, 0.4, 3000)
    pos["y"] = rng.normal(2.5, 0.4, 3000)
    with native.Renderer() as r:
        got = _gpu(r, pos, 6, 6, 10.0, abi.DEVIATION_FIT_UNCERTAINTY, [4.0, 1.0], abi.WEIGHING_INTEGRAL)
        empty = _gpu(r, pos[:0], 6, 6, 10.0, abi.DEVIATION_CONSTANT, [0.3], abi.WEIGHING_SAME)
    want = po.render(_as_rows(pos), 6, 6, 10.0, abi.DEVIATION_FIT_UNCERTAINTY, [4.0, 1.0], abi.WEIGHING_INTEGRAL)
    assert np.array_equal(got, want)
    assert empty.shape == (60, 60) and not empty.any()


def test_gpu_render_large_image():
    _say("case: render 40000 positions on 256x256 px x 8 (2048x2048 image)")
    rng = np.random.default_rng(11)
    pos = _positions(rng, 40000, 256, 256)
    with native.Renderer() as r:
        got = _gpu(r, pos, 256, 256, 8.0, abi.DEVIATION_CONSTANT, [0.2], abi.WEIGHING_INTEGRAL)
    want = po.render(_as_rows(pos), 256, 256, 8.0, abi.DEVIATION_CONSTANT, [0.2], abi.WEIGHING_INTEGRAL)
    assert np.array_equal(got, want)


def test_gpu_positions_from_fits():
    _say("case: accepted fits -> positions, in order")
    rng = np.random.default_rng(5)
    n = 5000
    fits = np.zeros(n, abi.FIT_DTYPE)
    fits["model"] = rng.choice([abi.MODEL_SYMMETRIC, abi.MODEL_ELLIPSOIDAL, abi.MODEL_FIXED_WIDTH], n)
    fits["status"] = np.where(rng.random(n) < 0.7, 0, rng.integers(1, 9, n))
    fits["params"] = rng.normal(size=(n, abi.MAX_PARAMS))
    fits["deviations"] = rng.random((n, abi.MAX_PARAMS))
    fits["integral"] = rng.uniform(1, 1000, n)
    d_fits = device.upload(fits)
    d_pos = device.empty(abi.POSITION_DTYPE, n)
    d_count = device.empty(np.int32, 1)
    with native.Renderer() as r:
        r.positions_from_fits_device(d_fits.ptr, n, d_pos.ptr, d_count.ptr, device.stream_handle())
    count = int(d_count.numpy()[0])
    got = d_pos.numpy()[:count]
    acc = fits[fits["status"] == 0]
    xi = np.where(acc["model"] == abi.MODEL_SYMMETRIC, 2, np.where(acc["model"] == abi.MODEL_FIXED_WIDTH, 1, 3))
    bi = np.where(acc["model"] == abi.MODEL_SYMMETRIC, 4, np.where(acc["model"] == abi.MODEL_FIXED_WIDTH, 3, 6))
    k = np.arange(len(acc))
    assert count == len(acc)
    assert np.array_equal(got["x"], acc["params"][k, xi])
    assert np.array_equal(got["y"], acc["params"][k, xi + 1])
    assert np.array_equal(got["integral"], acc["integral"])
    assert np.array_equal(got["x_deviation"], acc["deviations"][k, xi])
    assert np.array_equal(got["background"], acc["params"][k, bi])


def test_gpu_render_errors():
    _say("case: render error paths")
    pos = np.zeros(1, abi.POSITION_DTYPE)
    pos["integral"] = -1.0
    with native.Renderer() as r:
        with pytest.raises(native.LocalizerError, match="PSF width"):
            _gpu(r, pos, 8, 8, 2.0, abi.DEVIATION_GAUSSIAN_MASK, [1.3, 0.0, 1.0], abi.WEIGHING_SAME)
        with pytest.raises(native.LocalizerError, match="weighing"):
            d = device.upload(pos)
            img = device.empty(np.float64, 256)
            r.render_device(d.ptr, 1, 8, 8, 2.0, abi.DEVIATION_CONSTANT, [0.3], 5, img.ptr, device.stream_handle())
    with pytest.raises(native.LocalizerError):
        native.render_image_size(8, 8, 0.0)
