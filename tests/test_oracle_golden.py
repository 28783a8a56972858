"""The oracle's model and Jacobian against golden vectors written by the reference's
own functions (source/Localization.cpp:1007-1295, tests/golden/make_golden.py), and
the reference's Jacobian quirks against finite differences."""
import os

import numpy as np
import pytest

from oracle import pyoracle as po

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_models.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


def _cases(g):
    off_f = off_J = 0
    for c in range(len(g["kind"])):
        k, n = int(g["kind"][c]), int(g["n"][c])
        N, p = n * n, (5 if k == 0 else 7)
        yield (c, k, n, p, g["data"][off_f:off_f + N], g["f"][off_f:off_f + N],
               g["J"][off_J:off_J + N * p].reshape(N, p))
        off_f += N
        off_J += N * p


def test_model_and_jacobian_match_reference_functions(gold):
    checked = 0
    for c, k, n, p, data, fref, Jref in _cases(gold):
        th = gold["theta"][c][:p]
        st, f = po.model(k, th, data, n, gold["xoff"][c], gold["yoff"][c], gold["sigma"][c])
        stj, J = po.jacobian(k, th, n, gold["xoff"][c], gold["yoff"][c], gold["sigma"][c])
        assert st == gold["status_f"][c] and stj == gold["status_J"][c]
        if st != 0:
            assert st == -1  # GSL_FAILURE on zero width (Localization.cpp:1029-1031, :1098-1100)
            continue
        np.testing.assert_allclose(f, fref, rtol=1e-12, atol=1e-12 * np.abs(fref).max())
        scale = np.abs(Jref).max(axis=0)
        assert np.all(np.abs(J - Jref) <= 1e-12 * scale + 1e-300)
        checked += 1
    assert checked >= 50


def test_fixed_width_model_is_reference_symmetric_model_at_fixed_width(gold):
    """FitFunction_/Jacobian_FixedWidthGaussian (Localization.cpp:1050-1076, :1183-1225)
    evaluate the symmetric functions at r = width and keep columns A, x0, y0, b: the
    reference's own symmetric outputs pin them."""
    checked = 0
    for c, k, n, p, data, fref, Jref in _cases(gold):
        if k != 0 or gold["status_f"][c] != 0:
            continue
        th = gold["theta"][c][:5]
        t4 = th[[0, 2, 3, 4]]
        st, f = po.model_fixed_width(t4, th[1], data, n, gold["xoff"][c], gold["yoff"][c], gold["sigma"][c])
        stj, J = po.jacobian_fixed_width(t4, th[1], n, gold["xoff"][c], gold["yoff"][c], gold["sigma"][c])
        assert st == 0 and stj == 0
        np.testing.assert_allclose(f, fref, rtol=1e-12, atol=1e-12 * np.abs(fref).max())
        Jsel = Jref[:, [0, 2, 3, 4]]
        assert np.all(np.abs(J - Jsel) <= 1e-12 * np.abs(Jsel).max(axis=0) + 1e-300)
        checked += 1
    assert checked >= 20


def test_fixed_width_fit_recovers_noiseless_spot():
    """Known answer: a spot rendered exactly from the model at the fixed width is a zero
    of the residual; LM from a displaced start must land on it."""
    n, xoff, yoff, w = 9, 40.0, 70.0, 1.3
    truth = np.array([250.0, 44.3, 73.6, 20.0])
    data = po.model_fixed_width(truth, w, np.zeros(n * n), n, xoff, yoff)[1]  # model - 0
    t0 = truth + np.array([-60.0, 0.4, -0.35, 5.0])
    r = po.lm_run(3, data, n, xoff, yoff, t0, fixed_width=w)
    assert r["lm_status"] == 0
    np.testing.assert_allclose(r["theta"], truth, rtol=1e-7, atol=1e-6)


def test_eigen_closed_form_matches_eigen(gold):
    for (sx, sy, rho), info, ref in zip(gold["eig_in"], gold["eig_info"], gold["eig_out"]):
        st, sd1, sd2, theta = po.eigen2(sx, sy, rho)
        assert st == 0 and info == 0
        np.testing.assert_allclose([sd1, sd2], ref[:2], rtol=1e-13)
        # theta = atan(v0/v1); compare the slope (v0/v1) in a scale-safe way
        assert abs(np.sin(theta - ref[2])) < 1e-10


def test_eigen_rejects_non_finite():
    assert po.eigen2(np.nan, 1.0, 0.0)[0] == -1
    assert po.eigen2(1.0, np.inf, 0.0)[0] == -1


def _fd_jacobian(k, th, data, n, xoff, yoff, h=1e-6):
    p = th.size
    J = np.empty((n * n, p))
    for i in range(p):
        e = np.zeros(p)
        e[i] = h * max(1.0, abs(th[i]))
        J[:, i] = (po.model(k, th + e, data, n, xoff, yoff)[1] - po.model(k, th - e, data, n, xoff, yoff)[1]) / (2 * e[i])
    return J


def test_symmetric_width_column_is_true_derivative_over_sqrt2():
    """Quirk (Localization.cpp:1151-1152): d f/d r is reported as A e d^2/(sqrt(2) r^3)."""
    n, xoff, yoff = 9, 100.0, 200.0
    th = np.array([400.0, 1.4, 104.3, 203.8, 90.0])
    data = np.zeros(n * n)
    J = po.jacobian(0, th, n, xoff, yoff)[1]
    Jfd = _fd_jacobian(0, th, data, n, xoff, yoff)
    for col in (0, 2, 3, 4):
        np.testing.assert_allclose(J[:, col], Jfd[:, col], rtol=1e-6, atol=1e-6 * np.abs(Jfd[:, col]).max())
    np.testing.assert_allclose(J[:, 1] * np.sqrt(2.0), Jfd[:, 1], rtol=1e-6, atol=1e-6 * np.abs(Jfd[:, 1]).max())


def test_ellipsoidal_rho_column_exact_only_for_equal_widths():
    """Quirk (Localization.cpp:1280): the correlation column uses
    (dx^2 + dy^2 - 2 rho dx dy)/(sx sy) in place of the quadratic form."""
    n, xoff, yoff = 9, 10.0, 20.0
    data = np.zeros(n * n)
    th_eq = np.array([300.0, 1.5, 1.5, 14.2, 23.9, 0.3, 50.0])
    J = po.jacobian(1, th_eq, n, xoff, yoff)[1]
    Jfd = _fd_jacobian(1, th_eq, data, n, xoff, yoff)
    np.testing.assert_allclose(J, Jfd, rtol=1e-5, atol=1e-5 * np.abs(Jfd).max())
    th_ne = th_eq.copy()
    th_ne[2] = 2.0
    J = po.jacobian(1, th_ne, n, xoff, yoff)[1]
    Jfd = _fd_jacobian(1, th_ne, data, n, xoff, yoff)
    cols = [0, 1, 2, 3, 4, 6]
    np.testing.assert_allclose(J[:, cols], Jfd[:, cols], rtol=1e-5, atol=1e-5 * np.abs(Jfd).max())
    assert np.abs(J[:, 5] - Jfd[:, 5]).max() > 1e-2 * np.abs(Jfd[:, 5]).max()
