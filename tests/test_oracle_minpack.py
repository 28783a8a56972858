"""The oracle's LM engine (a restatement of GSL 1.16 lmsder, i.e. MINPACK lmder in
mode 1) against MINPACK itself: scipy.optimize.leastsq with a Jacobian calls lmder.
With the two GSL-specific choices switched to MINPACK behaviour (scale update over all
rows, no cap on retries) the accepted point after every trial evaluation must agree.
maxfev = k + 1 stops MINPACK after exactly k trial evaluations."""
import warnings

import numpy as np
import pytest
from scipy.optimize import leastsq

from localizer_amd import synth
from oracle import pyoracle as po


@pytest.fixture(scope="module")
def windows():
    w = synth.PLAIN_CPU_512.scaled(2)
    frames, _, cands = synth.generate(w)
    out = []
    for c in cands[:16]:
        h, n = 3, 7
        xoff, yoff = int(np.floor(c["x"] - h)), int(np.floor(c["y"] - h))
        data = frames[c["frame"], yoff:yoff + n, xoff:xoff + n].astype(float).reshape(-1)
        out.append((c, n, xoff, yoff, data))
    return out


@pytest.mark.parametrize("model", [0, 1, 3])
def test_trajectory_matches_minpack_lmder(windows, model):
    opt = po.default_options(minpack=True)
    points = 0
    w = 1.3
    for c, n, xoff, yoff, data in windows[:8]:
        if model == 0:
            t0 = np.array([c["intensity"], w, c["x"], c["y"], c["background"]])
        elif model == 3:
            t0 = np.array([c["intensity"], c["x"], c["y"], c["background"]])
        else:
            t0 = np.array([c["intensity"], w, w, c["x"], c["y"], 0.0, c["background"]])
        r = po.lm_run(model, data, n, xoff, yoff, t0, options=opt, max_trace=40, fixed_width=w)

        def fun(t):
            if model == 3:
                return po.model_fixed_width(t, w, data, n, xoff, yoff)[1]
            return po.model(model, t, data, n, xoff, yoff)[1]

        def jac(t):
            if model == 3:
                return po.jacobian_fixed_width(t, w, n, xoff, yoff)[1]
            return po.jacobian(model, t, n, xoff, yoff)[1]

        for k in range(1, min(r["trials"], 25) + 1):
            with warnings.catch_warnings():
                warnings.simplefilter("ignore")
                x, _ = leastsq(fun, t0, Dfun=jac, ftol=0.0, xtol=0.0, gtol=0.0, maxfev=k + 1)
            ref = r["trace"][k - 1]
            np.testing.assert_allclose(x, ref, rtol=1e-9, atol=1e-9)
            points += 1
    assert points >= (25 if model == 3 else 40)


def test_gsl_scale_update_changes_only_the_path(windows):
    """GSL 1.16's update_diag (first p rows) and MINPACK's (all rows) reach the same
    minimum; both stop by test_delta within its tolerance of each other."""
    for c, n, xoff, yoff, data in windows[:6]:
        t0 = np.array([c["intensity"], 1.3, c["x"], c["y"], c["background"]])
        a = po.lm_run(0, data, n, xoff, yoff, t0, options=po.default_options())
        b = po.lm_run(0, data, n, xoff, yoff, t0, options=po.default_options(minpack=True))
        assert a["lm_status"] == 0 and b["lm_status"] == 0
        tol = 1e-4 + 1e-4 * np.abs(a["theta"])
        assert np.all(np.abs(a["theta"] - b["theta"]) < 10 * tol)
