"""Candidate-generation oracle (oracle/lzr_find_oracle.c): the 8-way particle finder
against the reference's own ParticleFinder_adjacent8 (golden vectors,
tests/golden/make_golden_finder.py), and the GLRT segmentation against what the
mathematics fixes (the reference's GLRT needs FFTW, absent here)."""
import os

import numpy as np
import pytest

from oracle import pyoracle as po

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_finder.npz")


def test_particle_finder_matches_reference():
    g = np.load(GOLD)
    for k in range(int(g["n"][0])):
        got = po.find_particles(g[f"frame{k}"], g[f"mask{k}"])
        ref = g[f"particles{k}"]
        assert got.shape == ref.shape, k
        np.testing.assert_array_equal(got, ref)


def _gc(width):
    ws = po.glrt_window(width)
    h = ws // 2
    j, i = np.mgrid[-h:h + 1, -h:h + 1]
    g = 1.0 / (1.77245385 * width) * np.exp(-1.0 / (2 * width * width) * (i * i + j * j))
    return g - g.mean(), ws, h


@pytest.mark.parametrize("width,shape", [(1.3, (40, 52)), (1.0, (33, 47)), (2.2, (64, 64))])
def test_glrt_statistic_is_squared_correlation_with_the_spot(width, shape):
    """T = (sum gc x)^2 / (sum gc^2 * sum (x - mean)^2): the squared correlation of the
    window with the zero-mean Gaussian; tested range [h+1, X-h) x [h+1, Y-h) with X, Y
    the frame size rounded down to even (source/Segmentation.cpp:339-357, :418-425)."""
    rng = np.random.default_rng(7)
    frame = rng.integers(50, 400, shape).astype(np.uint16)
    mask, T = po.glrt_mask(frame, width, 20.0)
    gc, ws, h = _gc(width)
    ny, nx = shape
    X, Y = nx - nx % 2, ny - ny % 2
    tested = np.zeros(shape, bool)
    tested[h + 1:Y - h, h + 1:X - h] = True
    assert np.all(T[~tested] == 0) and np.all(mask[~tested] == 0)
    ys, xs = np.nonzero(tested)
    for y, x in zip(ys[::7], xs[::7]):
        win = frame[y - h:y + h + 1, x - h:x + h + 1].astype(np.float64)
        r2 = (gc * win).sum() ** 2 / ((gc * gc).sum() * ((win - win.mean()) ** 2).sum())
        assert abs(T[y, x] - r2) <= 1e-12 * max(1.0, r2)
    thr = 1 - np.exp(-20.0 / ws ** 2)
    np.testing.assert_array_equal(mask[tested].astype(bool), (T[tested] > thr) & (T[tested] < 1.0))


def test_reference_fft_path_equals_direct_correlation():
    """The reference convolves with the zero-mean Gaussian by FFT: kernel centred in a
    frame-sized image, product of the spectra times (-1)^(u+v) (source/Convolver.cpp:120-162,
    Segmentation.cpp:294-306).  Restated with numpy FFTs it equals the direct window
    correlation the oracle sums, wherever the window does not wrap."""
    rng = np.random.default_rng(3)
    ny, nx, width = 32, 48, 1.3
    frame = rng.integers(0, 1000, (ny, nx)).astype(np.float64)
    gc, ws, h = _gc(width)
    big = np.zeros((ny, nx))
    cy, cx = ny // 2, nx // 2
    big[cy - h:cy + h + 1, cx - h:cx + h + 1] = gc
    F = np.fft.rfft2(frame) * np.fft.rfft2(big)
    v, u = np.meshgrid(np.arange(F.shape[0]), np.arange(F.shape[1]), indexing="ij")
    conv = np.fft.irfft2(F * np.where((u + v) % 2 == 0, 1.0, -1.0), s=(ny, nx))
    for y in range(h + 1, ny - h):
        for x in range(h + 1, nx - h):
            direct = (frame[y - h:y + h + 1, x - h:x + h + 1] * gc).sum()
            assert abs(conv[y, x] - direct) <= 1e-9 * np.abs(frame).sum()


def test_glrt_detects_a_spot_and_ignores_flat_frames():
    ny, nx, width = 40, 40, 1.3
    flat = np.full((ny, nx), 100, np.uint16)
    mask, T = po.glrt_mask(flat, width, 20.0)
    assert mask.sum() == 0  # zero variance: T is NaN, never segmented
    y, x = np.mgrid[0:ny, 0:nx]
    spot = flat + np.round(800 * np.exp(-((x - 20.3) ** 2 + (y - 19.6) ** 2) / (2 * width ** 2))).astype(np.uint16)
    mask, T = po.glrt_mask(spot, width, 20.0)
    assert mask[20, 20] == 1 and mask[5, 5] == 0
    parts = po.find_particles(spot, mask)
    assert len(parts) == 1
    assert abs(parts[0, 0] - 20.3) < 0.6 and abs(parts[0, 1] - 19.6) < 0.6
    bg = spot[mask == 0].astype(np.float64).mean()  # mean of every unsegmented pixel
    assert parts[0, 3] == bg and parts[0, 2] == spot.max() - bg
