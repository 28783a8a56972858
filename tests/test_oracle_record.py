"""Pins of the oracle's record side (oracle/lzr_oracle.c) by what the mathematics fixes.

* orc_covar is gsl_multifit_covar(J, 0.0, C) (source/Localization.cpp:185, :556): the
  inverse of J^T J, checked against numpy.linalg.inv on scaled random Jacobians; with
  epsrel = 0 an exactly-zero column is dropped (its rows and columns of C are 0, the rest is
  the inverse over the remaining columns).
* The integral (source/Localization.cpp:202, :371, :582-583) against numerical quadrature of
  the fitted surface minus the offset: 2 pi r^2 A for the symmetric and fixed-width models;
  for the ellipsoidal model the reference's 2 pi A sqrt(1 - rho^2) sd1 sd2 is the volume
  times a further sqrt(1 - rho^2), since sd1 sd2 = sx sy sqrt(1 - rho^2) already (a quirk
  kept for parity, exact when rho = 0).
* A noise-free spot's fitted integral is its photon count (A = N / (2 pi sigma^2)).
The formulas of integralDeviation and of the ellipsoidal sd-deviation mixing
(:206-208, :371-372, :575-592) are restated from the reference and not fixed by mathematics:
parity unpinned beyond reading."""
import math

import numpy as np
import pytest

from localizer_amd import abi
from oracle import pyoracle as po


@pytest.mark.parametrize("p", [4, 5, 7])
def test_covar_is_inverse_of_normal_matrix(p):
    rng = np.random.default_rng(p)
    for _ in range(20):
        J = rng.normal(size=(81, p)) * 10.0 ** rng.uniform(-3, 3, p)  # columns of very different scale
        C = po.covar(J)
        want = np.linalg.inv(J.T @ J)
        d = np.sqrt(np.diag(want))
        assert np.allclose(C / np.outer(d, d), want / np.outer(d, d), rtol=0, atol=1e-10)


def test_covar_drops_exactly_zero_columns():
    rng = np.random.default_rng(1)
    J = rng.normal(size=(49, 5))
    J[:, 2] = 0.0
    C = po.covar(J)
    keep = [0, 1, 3, 4]
    assert np.all(C[2, :] == 0) and np.all(C[:, 2] == 0)
    assert np.allclose(C[np.ix_(keep, keep)], np.linalg.inv(J[:, keep].T @ J[:, keep]), rtol=1e-12, atol=0)


def _spot(n, A, sx, sy, x0, y0, rho, b):
    jj, ii = np.mgrid[0:n, 0:n].astype(np.float64)
    u, v = ii - x0, jj - y0
    q = u * u / sx ** 2 + v * v / sy ** 2 - 2 * rho * u * v / (sx * sy)
    return b + A * np.exp(-q / (2 * (1 - rho * rho)))


def _fit_one(model, data, x, y, A0, b0, w, h):
    n = 2 * h + 1
    frame = np.zeros((1, 64, 64), np.float64)
    frame[0, 20:20 + n, 20:20 + n] = data
    c = np.zeros(1, abi.CANDIDATE_DTYPE)
    c["x"], c["y"], c["intensity"], c["background"] = 20 + x, 20 + y, A0, b0
    c["half_width"] = h
    c["model"] = -1
    # the oracle takes uint16 frames: noise-free values are rounded, so use large amplitudes
    r = po.fit_batch(np.round(frame).astype(np.uint16), c, model, w, default_half_width=h)[0]
    assert r["status"] == abi.ACCEPTED
    return r


def _quadrature(f, half=12.0, k=2401):
    t = np.linspace(-half, half, k)
    X, Y = np.meshgrid(t, t, indexing="ij")
    z = f(X, Y)
    return np.trapezoid(np.trapezoid(z, t, axis=1), t)


def test_symmetric_integral_is_volume_and_photon_count():
    N, sigma, b = 400000.0, 1.4, 100.0
    A = N / (2 * math.pi * sigma ** 2)
    data = _spot(15, A, sigma, sigma, 7.3, 6.8, 0.0, b)
    r = _fit_one(abi.MODEL_SYMMETRIC, data, 7, 7, A * 0.8, b, 1.35, 7)
    A1, r1 = r["params"][0], r["params"][1]
    vol = _quadrature(lambda X, Y: A1 * np.exp(-(X * X + Y * Y) / (2 * r1 * r1)))
    assert r["integral"] == pytest.approx(vol, rel=1e-9)
    assert r["integral"] == pytest.approx(N, rel=2e-4)  # uint16 rounding of the data


def test_fixed_width_integral_is_volume():
    w, b = 1.35, 100.0
    A = 30000.0
    data = _spot(13, A, w, w, 6.2, 5.7, 0.0, b)
    r = _fit_one(abi.MODEL_FIXED_WIDTH, data, 6, 6, A * 0.9, b, w, 6)
    A1 = r["params"][0]
    vol = _quadrature(lambda X, Y: A1 * np.exp(-(X * X + Y * Y) / (2 * w * w)))
    assert r["integral"] == pytest.approx(vol, rel=1e-9)


@pytest.mark.parametrize("rho", [0.0, 0.45])
def test_ellipsoidal_integral_is_volume_times_sqrt_one_minus_rho2(rho):
    sx, sy, b = 1.6, 1.2, 100.0
    A = 30000.0
    data = _spot(17, A, sx, sy, 8.2, 7.9, rho, b)
    r = _fit_one(abi.MODEL_ELLIPSOIDAL, data, 8, 8, A * 0.9, b, 1.35, 8)
    A1, sx1, sy1, rho1 = r["params"][0], r["params"][1], r["params"][2], r["params"][5]
    assert rho1 == pytest.approx(rho, abs=1e-3)

    def surface(X, Y):
        q = X * X / sx1 ** 2 + Y * Y / sy1 ** 2 - 2 * rho1 * X * Y / (sx1 * sy1)
        return A1 * np.exp(-q / (2 * (1 - rho1 * rho1)))

    vol = _quadrature(surface, half=14.0, k=2801)
    assert r["integral"] == pytest.approx(vol * math.sqrt(1 - rho1 * rho1), rel=1e-8)
    # sd1 sd2 is the square root of the covariance determinant
    assert r["sd1"] * r["sd2"] == pytest.approx(sx1 * sy1 * math.sqrt(1 - rho1 * rho1), rel=1e-12)


def test_fp64_frames_of_counts_fit_like_uint16_frames():
    """orc_fit_batch_f64 reads the frame as the reference's FP64 Image: integer-valued frames
    give the uint16 driver's records bit for bit; a constant offset added to every pixel moves
    only the fitted background, by that offset (to the LM's tolerance)."""
    from localizer_amd import synth

    frames, _, cands = synth.generate(synth.PLAIN_CPU_512.scaled(4))
    a = po.fit_batch(frames, cands, abi.MODEL_SYMMETRIC, 1.3)
    b = po.fit_batch(frames.astype(np.float64), cands, abi.MODEL_SYMMETRIC, 1.3)
    assert a.tobytes() == b.tobytes()
    c = po.fit_batch(frames.astype(np.float64) + 0.375, cands, abi.MODEL_SYMMETRIC, 1.3)
    both = (a["status"] == 0) & (c["status"] == 0)
    assert both.sum() > 0.9 * (a["status"] == 0).sum()
    assert np.allclose(c["params"][both, 4] - a["params"][both, 4], 0.375, atol=2e-3)
    assert np.allclose(c["params"][both, 2:4], a["params"][both, 2:4], atol=1e-3)
