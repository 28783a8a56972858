"""Convergence of the reference LM from near-optimal starts (oracle only).

The reference's symmetric width column is the true derivative divided by sqrt(2)
(source/Localization.cpp:1151-1152), so Gauss-Newton steps overshoot the width and the
iteration converges linearly: the number of trials grows with the logarithm of the
start error.  This bounds what any start-point predictor can save."""
import numpy as np

from localizer_amd import abi, netinit, synth
from oracle import pyoracle as po

W = synth.LEARNED_INIT_11


def test_trials_grow_with_start_error_even_near_the_solution():
    rois, _, cands = netinit.simulate(W, 300, 93)
    ref = po.fit_batch(rois, cands, abi.MODEL_SYMMETRIC, W.initial_psf_width, default_half_width=5)
    idx = np.nonzero((ref["status"] == 0) & (ref["lm_status"] == 0))[0][:60]
    rng = np.random.default_rng(0)
    trials = {}
    for noise in (0.0, 0.01, 0.05):
        t = []
        for i in idx:
            p = ref["params"][i][:5].copy()
            t0 = p.copy()
            t0[2:4] += rng.normal(0, noise, 2)
            t0[1] *= np.exp(rng.normal(0, noise / 2))
            t0[0] *= np.exp(rng.normal(0, noise))
            r = po.lm_run(abi.MODEL_SYMMETRIC, rois[i].astype(float).reshape(-1), 11, 0.0, 0.0, t0)
            t.append(r["trials"])
        trials[noise] = float(np.mean(t))
    plain = float(ref["trials"][idx].mean())
    print({"plain": plain, **{f"start error {k}": v for k, v in trials.items()}})
    assert trials[0.0] <= 2.0                      # at the solution: one step
    assert trials[0.01] >= 4.0                     # 0.01 px away already needs several
    assert trials[0.05] >= 0.45 * plain            # 0.05 px away: about half the plain start's
    assert trials[0.0] < trials[0.01] < trials[0.05] < plain
