"""The separable-moment pass of the symmetric and fixed-width models
(localizer_amd/csrc/lzr_lane.h sep_pass / sep_finish, host build) against brute-force
per-pixel FP64 sums.

The pass never forms the Jacobian: it combines 1-D moments of the Gaussian factors with
five data sums.  Here the same quantities are summed pixel by pixel from the reference's
model and Jacobian columns (source/Localization.cpp:1007-1048, :1128-1181; the width
column is the true derivative / sqrt(2)), at trial points that exercise the factor
recurrence, its direct-exponential fallback (collapsed widths, centres far outside the
window) and the masking of a lane's window inside a larger loop bound."""
import math

import numpy as np
import pytest

from tests.emu import pyemu


def _brute(theta, dx, sigma, w, window):
    t = np.asarray(theta, np.float64) + np.asarray(dx, np.float64)
    p = t.size
    n = window.shape[0]
    jj, ii = np.mgrid[0:n, 0:n].astype(np.float64)
    D = window.astype(np.float64)
    if p == 5:
        A, r, x0, y0, b = t
    else:
        A, x0, y0, b = t
        r = w
    u, v = ii - x0, jj - y0
    d2 = u * u + v * v
    e = np.exp(-d2 / (2 * r * r))
    f = (A * e + b - D) / sigma
    cols = [e / sigma]
    if p == 5:
        cols.append(A * e * d2 / (math.sqrt(2.0) * r ** 3 * sigma))
    cols += [A * e * u / (r * r * sigma), A * e * v / (r * r * sigma)]
    J = np.stack([c.ravel() for c in cols], axis=1)
    H = J.T @ J
    hq = np.array([H[i, j] for i in range(p - 1) for j in range(i + 1)])
    want = (hq, J.sum(axis=0), J.T @ f.ravel(), f.sum(), float(f.ravel() @ f.ravel()))
    # magnitudes of the terms each sum cancels (model against data): the FP64 floor
    fa = ((np.abs(A * e) + abs(b) + D) / sigma).ravel()
    aJ = np.abs(J)
    mag = (np.array([(aJ[:, i] * aJ[:, j]).sum() for i in range(p - 1) for j in range(i + 1)]), aJ.sum(axis=0),
           aJ.T @ fa, fa.sum(), float(fa @ fa))
    return want, mag


def _check(theta, dx, sigma, w, window, nloop=None):
    got = pyemu.sep_pass(theta, dx, sigma, w, window, nloop)
    want, mag = _brute(theta, dx, sigma, w, window)
    names = ("J^T J", "J^T 1", "J^T f", "sum f", "|f|^2")
    for name, g, r, m in zip(names, got, want, mag):
        g, r, m = np.atleast_1d(g), np.atleast_1d(r), np.atleast_1d(m)
        assert np.all(np.isfinite(g)), name
        # within 1e-12 of the summed term magnitudes (and 1e-11 relative where nothing cancels)
        assert np.all(np.abs(g - r) <= 1e-11 * np.abs(r) + 1e-12 * m + 1e-300), (name, g, r, m)


def _window(rng, n, A, r, x0, y0, b):
    jj, ii = np.mgrid[0:n, 0:n]
    lam = b + A * np.exp(-((ii - x0) ** 2 + (jj - y0) ** 2) / (2 * r * r))
    return rng.poisson(lam).astype(np.uint16)


@pytest.mark.parametrize("n", [5, 7, 9, 15])
def test_symmetric_pass_matches_pixel_sums(n):
    rng = np.random.default_rng(n)
    for _ in range(25):
        c = (n - 1) / 2
        theta = [rng.uniform(100, 3000), rng.uniform(0.8, 2.5), c + rng.uniform(-1.5, 1.5), c + rng.uniform(-1.5, 1.5),
                 rng.uniform(50, 200)]
        dx = rng.normal(0, [30, 0.05, 0.2, 0.2, 3])
        window = _window(rng, n, theta[0], theta[1], theta[2], theta[3], theta[4])
        _check(theta, dx, 1.0, 1.3, window)
        _check(theta, np.zeros(5), 1.0, 1.3, window)  # the start point (dx = 0)


def test_fixed_width_pass_matches_pixel_sums():
    rng = np.random.default_rng(4)
    for n, w in ((5, 1.0), (9, 1.35), (11, 1.8)):
        c = (n - 1) / 2
        for _ in range(20):
            theta = [rng.uniform(100, 3000), c + rng.uniform(-2, 2), c + rng.uniform(-2, 2), rng.uniform(50, 200)]
            dx = rng.normal(0, [30, 0.2, 0.2, 3])
            _check(theta, dx, 1.0, w, _window(rng, n, theta[0], w, theta[1], theta[2], theta[3]))


def test_recurrence_fallback_and_extreme_trial_points():
    """Exponents beyond the recurrence's safe range (a collapsed width: c = 1/(2 r^2) up to
    5e3; centres far outside the window) take the direct exponentials; very wide spots
    stay on the recurrence.  A residual weight sigma != 1 scales every column."""
    rng = np.random.default_rng(9)
    window = _window(rng, 9, 800, 1.3, 4.2, 3.7, 100)
    for theta in ([800, 0.01, 4.2, 3.7, 100],   # c = 5000: direct factors, most underflow
                  [800, 0.08, 4.5, 4.5, 100],   # c ~ 78: direct
                  [800, 0.3, -6.0, 4.0, 100],   # centre 6 px left of the window
                  [800, 1.3, 4.0, 30.0, 100],   # centre far below
                  [800, 40.0, 4.0, 4.0, 100],   # nearly flat
                  [800, 1.3, 4.0, 4.0, 100]):
        _check(theta, np.zeros(5), 1.0, 1.3, window)
        _check(theta, np.zeros(5), 2.5, 1.3, window)


def test_masked_window_inside_a_larger_loop():
    """A lane with an n x n window in a wave whose loop runs to nloop > n: columns and rows
    beyond n must contribute nothing (the GPU masks them the same way)."""
    rng = np.random.default_rng(3)
    for n, nloop in ((5, 9), (7, 15), (9, 11)):
        c = (n - 1) / 2
        theta = [1500.0, 1.2, c + 0.3, c - 0.4, 120.0]
        window = _window(rng, n, *theta)
        full = pyemu.sep_pass(theta, np.zeros(5), 1.0, 1.3, window, n)
        masked = pyemu.sep_pass(theta, np.zeros(5), 1.0, 1.3, window, nloop)
        for a, b in zip(full, masked):
            np.testing.assert_allclose(a, b, rtol=1e-13, atol=1e-9)
        _check(theta, np.zeros(5), 1.0, 1.3, window, nloop)


def _brute_ell(theta, dx, sigma, window):
    """Per-pixel FP64 sums from the reference's ellipsoidal model and Jacobian columns
    (source/Localization.cpp:1078-1126, :1227-1295; the rho column's quirk included)."""
    A, sx, sy, x0, y0, rho, b = np.asarray(theta, np.float64) + np.asarray(dx, np.float64)
    n = window.shape[0]
    jj, ii = np.mgrid[0:n, 0:n].astype(np.float64)
    D = window.astype(np.float64)
    u, v = ii - x0, jj - y0
    omr = 1.0 - rho * rho
    Q = u * u / sx ** 2 + v * v / sy ** 2 - 2 * rho * u * v / (sx * sy)
    e = np.exp(-Q / (2 * omr))
    f = (A * e + b - D) / sigma
    den = 2 * omr * sigma
    cols = [e / sigma,
            -A * (2 * rho * u * v / (sx * sx * sy) - 2 * u * u / sx ** 3) * e / den,
            -A * (2 * rho * u * v / (sx * sy * sy) - 2 * v * v / sy ** 3) * e / den,
            -A * (2 * rho * v / (sx * sy) - 2 * u / (sx * sx)) * e / den,
            -A * (2 * rho * u / (sx * sy) - 2 * v / (sy * sy)) * e / den,
            A * (u * v / (omr * sx * sy) - rho * (u * u + v * v - 2 * rho * u * v) / (sx * sy * omr * omr)) * e / sigma]
    J = np.stack([c.ravel() for c in cols], axis=1)
    H = J.T @ J
    hq = np.array([H[i, j] for i in range(6) for j in range(i + 1)])
    want = (hq, J.sum(axis=0), J.T @ f.ravel(), f.sum(), float(f.ravel() @ f.ravel()))
    fa = ((np.abs(A * e) + abs(b) + D) / sigma).ravel()
    aJ = np.abs(J)
    mag = (np.array([(aJ[:, i] * aJ[:, j]).sum() for i in range(6) for j in range(i + 1)]), aJ.sum(axis=0),
           aJ.T @ fa, fa.sum(), float(fa @ fa))
    return want, mag


def _check_ell(theta, dx, sigma, window, nloop=None, rtol=1e-11, mtol=1e-12):
    got = pyemu.ell_pass(theta, dx, sigma, window, nloop)
    want, mag = _brute_ell(theta, dx, sigma, window)
    for name, g, r, m in zip(("J^T J", "J^T 1", "J^T f", "sum f", "|f|^2"), got, want, mag):
        g, r, m = np.atleast_1d(g), np.atleast_1d(r), np.atleast_1d(m)
        assert np.all(np.isfinite(g)), name
        assert np.all(np.abs(g - r) <= rtol * np.abs(r) + mtol * m + 1e-300), (name, g, r, m)


def _window_ell(rng, n, A, sx, sy, x0, y0, rho, b):
    jj, ii = np.mgrid[0:n, 0:n]
    u, v = ii - x0, jj - y0
    Q = u * u / sx ** 2 + v * v / sy ** 2 - 2 * rho * u * v / (sx * sy)
    return rng.poisson(b + A * np.exp(-Q / (2 * (1 - rho * rho)))).astype(np.uint16)


@pytest.mark.parametrize("n", [5, 7, 9, 11, 15])
def test_ellipsoidal_moment_pass_matches_pixel_sums(n):
    """The ellipsoidal moment pass (Ex Ey C factorisation, the cross factor by recurrence
    along rows) against the reference's columns summed pixel by pixel."""
    rng = np.random.default_rng(20 + n)
    c = (n - 1) / 2
    for _ in range(25):
        theta = [rng.uniform(100, 3000), rng.uniform(0.8, 2.2), rng.uniform(0.8, 2.2), c + rng.uniform(-1.5, 1.5),
                 c + rng.uniform(-1.5, 1.5), rng.uniform(-0.6, 0.6), rng.uniform(50, 200)]
        dx = rng.normal(0, [30, 0.05, 0.05, 0.2, 0.2, 0.05, 3])
        window = _window_ell(rng, n, *theta)
        _check_ell(theta, dx, 1.0, window)
        _check_ell(theta, np.zeros(7), 1.0, window)


def test_ellipsoidal_direct_fallback_and_masking():
    """Trial points whose factors would overflow on their own (|rho| near 1, collapsed
    widths) take per-pixel exponentials; a lane's window inside a larger loop bound
    contributes nothing beyond n."""
    rng = np.random.default_rng(31)
    window = _window_ell(rng, 11, 900, 1.4, 1.1, 5.2, 4.6, 0.2, 100)
    for theta in ([900, 1.4, 1.1, 5.2, 4.6, 0.2, 100],
                  [900, 1.4, 1.1, 5.2, 4.6, 0.995, 100],   # 1/(1-rho^2) ~ 100: direct
                  [900, 0.05, 1.1, 5.2, 4.6, 0.2, 100],    # collapsed sx: direct
                  [900, 1.4, 1.1, -4.0, 4.6, -0.5, 100],   # centre outside the window
                  [900, 30.0, 25.0, 5.0, 5.0, 0.0, 100]):  # nearly flat
        _check_ell(theta, np.zeros(7), 1.0, window)
        _check_ell(theta, np.zeros(7), 2.0, window, nloop=15)
        _check_ell(theta, np.zeros(7), 2.0, window, nloop=13)

    # strong coupling (|qc| ~ 6.6): factored inside the 5 x 5 window, but the cross factor
    # overflows in the loop's columns and rows beyond it; those must contribute exactly 0
    small = _window_ell(rng, 5, 900, 0.6, 0.6, 2.0, 2.0, 0.9, 100)
    theta = [900, 0.6, 0.6, 2.0, 2.0, 0.9, 100]
    _check_ell(theta, np.zeros(7), 1.0, small, nloop=15, rtol=1e-10, mtol=1e-11)
    for a, b in zip(pyemu.ell_pass(theta, np.zeros(7), 1.0, small, 5), pyemu.ell_pass(theta, np.zeros(7), 1.0, small, 15)):
        assert np.array_equal(np.atleast_1d(a), np.atleast_1d(b))


@pytest.mark.parametrize("n", [5, 9, 11, 15])
def test_group_pass_bit_identical_to_one_lane(n):
    """A straggler group of k lanes splits a window's rows over its lanes and folds their
    results in row order after an exchange (lzr_fit.hip straggler_groups).  Emulated lane by
    lane on the host, every k in 1..16 must give the one-lane pass's sums bit for bit, for the
    separable and the ellipsoidal passes, including the direct-exponential fallbacks and a
    residual overflow (|rho| > 1) whose count the ellipsoidal pass carries."""
    rng = np.random.default_rng(40 + n)
    c = (n - 1) / 2
    for trial in range(12):
        t5 = [rng.uniform(20, 3000), rng.uniform(0.6, 2.2), c + rng.uniform(-1.5, 1.5), c + rng.uniform(-1.5, 1.5),
              rng.uniform(50, 200)]
        if trial == 0:
            t5[1] = 0.03  # collapsed width: per-column exponentials
        win = _window(rng, n, *t5)
        dx5 = rng.normal(0, [30, 0.05, 0.2, 0.2, 3])
        one = pyemu.sep_pass_flat(t5, dx5, 1.0, 1.35, win)
        t4 = [t5[0], t5[2], t5[3], t5[4]]
        one4 = pyemu.sep_pass_flat(t4, dx5[[0, 2, 3, 4]], 1.0, 1.35, win)
        t7 = [rng.uniform(100, 3000), rng.uniform(0.8, 2.2), rng.uniform(0.8, 2.2), c + rng.uniform(-1.5, 1.5),
              c + rng.uniform(-1.5, 1.5), rng.uniform(-0.6, 0.6), rng.uniform(50, 200)]
        if trial == 1:
            t7[5] = 0.995  # unfactored
        if trial == 2:
            t7[5] = -1.02  # |rho| > 1: overflowing residuals
        wel = _window_ell(rng, n, *[t7[0], t7[1], t7[2], t7[3], t7[4], min(abs(t7[5]), 0.9), t7[6]])
        dx7 = rng.normal(0, [30, 0.05, 0.05, 0.2, 0.2, 0.01, 3])
        one7 = pyemu.ell_pass_flat(t7, dx7, 1.0, wel)
        for k in (1, 2, 4, 8, 16):
            assert np.array_equal(pyemu.sep_pass_group(t5, dx5, 1.0, 1.35, win, k), one, equal_nan=True), (trial, k)
            assert np.array_equal(pyemu.sep_pass_group(t4, dx5[[0, 2, 3, 4]], 1.0, 1.35, win, k), one4,
                                  equal_nan=True), (trial, k)
            assert np.array_equal(pyemu.ell_pass_group(t7, dx7, 1.0, wel, k), one7, equal_nan=True), (trial, k)
