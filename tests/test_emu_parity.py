"""The GPU lane logic (localizer_amd/csrc/lzr_lane.h), compiled for the host, against the
FP64 oracle.  This checks the normal-equation restatement of lmsder, the moment passes and
the record epilogue on the CPU, to the same bar as the GPU tests (tests/parity.py); the same
code runs on the GPU in tests/test_gpu_parity.py."""
import numpy as np
import pytest

from localizer_amd import abi, synth
from oracle import pyoracle as po
from tests.emu import pyemu
from tests.parity import assert_parity, compare, format_stats, oracle_sigma_ulp


@pytest.fixture(scope="module")
def plain_cpu():
    return synth.generate(synth.PLAIN_CPU_512)


@pytest.fixture(scope="module")
def plain_gpu_sample():
    return synth.generate(synth.PLAIN_GPU_256.scaled(60))


def test_emulated_lanes_match_oracle_7x7(plain_cpu):
    frames, _, cands = plain_cpu
    ref = po.fit_batch(frames, cands, abi.MODEL_SYMMETRIC, 1.3, default_half_width=3)
    got = pyemu.fit_batch(frames, cands, abi.MODEL_SYMMETRIC, 1.3, default_half_width=3)
    st = compare(ref, got, abi.MODEL_SYMMETRIC)
    print(format_stats("emu 7x7", st))
    assert_parity(st, abi.MODEL_SYMMETRIC)
    assert st["xy_frac_within_1e-3"] == 1.0


def test_emulated_lanes_match_oracle_9x9(plain_gpu_sample):
    frames, _, cands = plain_gpu_sample
    ref = po.fit_batch(frames, cands, abi.MODEL_SYMMETRIC, 1.35, default_half_width=4)
    got = pyemu.fit_batch(frames, cands, abi.MODEL_SYMMETRIC, 1.35, default_half_width=4)
    st = compare(ref, got, abi.MODEL_SYMMETRIC)
    print(format_stats("emu 9x9", st))
    assert_parity(st, abi.MODEL_SYMMETRIC)


@pytest.fixture(scope="module")
def plain_20k():
    return synth.generate(synth.PLAIN_GPU_256.scaled(400), nthreads=8)


def test_emulated_lanes_match_oracle_ellipsoidal(plain_20k):
    """20k fits: the 0.1 % bar is then a rate, not a matter of two or three fits."""
    frames, _, cands = plain_20k
    ref = po.fit_batch(frames, cands, abi.MODEL_ELLIPSOIDAL, 1.35, default_half_width=4)
    got = pyemu.fit_batch(frames, cands, abi.MODEL_ELLIPSOIDAL, 1.35, default_half_width=4)
    st = compare(ref, got, abi.MODEL_ELLIPSOIDAL)
    print(format_stats("emu ellipsoidal 9x9", st))
    spread = compare(ref, oracle_sigma_ulp(po, frames, cands, abi.MODEL_ELLIPSOIDAL, 1.35, default_half_width=4),
                     abi.MODEL_ELLIPSOIDAL)
    assert_parity(st, abi.MODEL_ELLIPSOIDAL, spread=spread)


def test_emulated_lanes_match_oracle_fixed_width(plain_gpu_sample):
    frames, _, cands = plain_gpu_sample
    ref = po.fit_batch(frames, cands, abi.MODEL_FIXED_WIDTH, 1.35, default_half_width=4)
    got = pyemu.fit_batch(frames, cands, abi.MODEL_FIXED_WIDTH, 1.35, default_half_width=4)
    st = compare(ref, got, abi.MODEL_FIXED_WIDTH)
    print(format_stats("emu fixed width 9x9", st))
    assert_parity(st, abi.MODEL_FIXED_WIDTH)
    acc = (ref["status"] == 0) & (got["status"] == 0)
    assert np.all(got["sd1"][acc] == 1.35)


def test_fixed_width_distance_rejection():
    """A candidate more than 2 w from the spot it converges to is rejected
    (Localization.cpp:349-353); the symmetric fitter has no such check."""
    frames = np.full((1, 48, 48), 100, np.uint16)
    yy, xx = np.mgrid[0:48, 0:48]
    frames[0] += np.round(500.0 * np.exp(-((xx - 20.0) ** 2 + (yy - 20.0) ** 2) / (2 * 1.0 ** 2))).astype(np.uint16)
    c = np.zeros(2, abi.CANDIDATE_DTYPE)
    c["x"] = [22.8, 20.4]
    c["y"] = [20.0, 20.2]
    c["intensity"] = 400
    c["background"] = 100
    c["half_width"] = 4
    c["model"] = -1
    for fit in (lambda m: po.fit_batch(frames, c, m, 1.0, default_half_width=4),
                lambda m: pyemu.fit_batch(frames, c, m, 1.0, default_half_width=4)):
        fw = fit(abi.MODEL_FIXED_WIDTH)
        assert list(fw["status"]) == [abi.REJECT_DISTANCE, abi.ACCEPTED]
        assert abs(fw["params"][1][1] - 20.0) < 0.01
        sym = fit(abi.MODEL_SYMMETRIC)
        assert abi.REJECT_DISTANCE not in sym["status"]


def test_nonfinite_trial_residual_follows_gsl_norm(plain_20k):
    """A trial step to |rho| > 1 overflows the ellipsoidal model: the reference's residual
    vector then holds infinite entries and GSL's scaled 2-norm is NaN for two or more of them
    (+inf for one), which sends lmsder's trust-region update down a different branch than a
    merely large norm.  Spots whose reference trajectory passes through a NaN-norm trial must
    follow the oracle (same iteration and trial counts, position within 1e-3 px)."""
    frames, _, cands = plain_20k
    sel = np.arange(0, cands.size, 4)
    ref = po.fit_batch(frames, cands[sel], abi.MODEL_ELLIPSOIDAL, 1.35, default_half_width=4)
    got = pyemu.fit_batch(frames, cands[sel], abi.MODEL_ELLIPSOIDAL, 1.35, default_half_width=4)
    hit = agree = 0
    for k in np.nonzero(ref["status"] == 0)[0]:
        c = cands[sel[k]]
        xo, yo = int(np.floor(c["x"] - 4)), int(np.floor(c["y"] - 4))
        D = frames[c["frame"], yo:yo + 9, xo:xo + 9].astype(np.float64)
        r = po.lm_run(abi.MODEL_ELLIPSOIDAL, D, 9, xo, yo,
                      np.array([c["intensity"], 1.35, 1.35, c["x"], c["y"], 0.0, c["background"]]), max_trace=250)
        if not np.isnan(r["trial_norms"]).any():
            continue
        hit += 1
        agree += (got["trials"][k] == ref["trials"][k] and got["iterations"][k] == ref["iterations"][k]
                  and np.abs(got["params"][k][3:5] - ref["params"][k][3:5]).max() <= 1e-3)
    print(f"{hit} spots with a NaN-norm trial, {agree} follow the oracle")
    assert hit >= 3
    assert agree == hit


@pytest.mark.parametrize("model", [abi.MODEL_SYMMETRIC, abi.MODEL_FIXED_WIDTH, abi.MODEL_ELLIPSOIDAL])
def test_start_residual_matches_pixel_sum(plain_gpu_sample, model):
    """The lane's residual at the reference start point (the net-start fallback's yardstick,
    from the moment pass) equals the oracle's |f| summed pixel by pixel at that point
    (source/Localization.cpp:138-142: intensity, initial width, position, background; rho 0),
    to 1e-10 relative; windows outside the frame give NaN."""
    frames, _, cands = plain_gpu_sample
    cands = cands[:3000].copy()
    # the default window, and every window class of the unrolled loops (n = 5 ... 31)
    cands["half_width"] = np.random.default_rng(5).choice([0, 2, 5, 6, 9, 10, 15], len(cands))
    cands["x"][:3] = [-4.0, 2.0, frames.shape[2] + 1.0]  # outside on the left / near edge / right
    cands["half_width"][:3] = 0
    w = synth.PLAIN_GPU_256.initial_psf_width
    got = pyemu.start_residual(frames, cands, model, w)
    ny, nx = frames.shape[1:]
    want = np.full(len(cands), np.nan)
    for i, c in enumerate(cands):
        h = int(c["half_width"]) or abi.default_half_width(w)
        n = 2 * h + 1
        xoff, yoff = np.floor(c["x"] - h), np.floor(c["y"] - h)
        if xoff < 0 or yoff < 0 or xoff + n > nx or yoff + n > ny:
            continue
        data = frames[c["frame"], int(yoff):int(yoff) + n, int(xoff):int(xoff) + n]
        if model == abi.MODEL_SYMMETRIC:
            _, f = po.model(model, [c["intensity"], w, c["x"], c["y"], c["background"]], data, n, xoff, yoff)
        elif model == abi.MODEL_FIXED_WIDTH:
            _, f = po.model_fixed_width([c["intensity"], c["x"], c["y"], c["background"]], w, data, n, xoff, yoff)
        else:
            _, f = po.model(model, [c["intensity"], w, w, c["x"], c["y"], 0.0, c["background"]], data, n, xoff, yoff)
        want[i] = np.sqrt(np.sum(f * f))
    assert np.array_equal(np.isnan(got), np.isnan(want)) and np.isnan(want[[0, 2]]).all()
    ok = ~np.isnan(want)
    rel = np.abs(got[ok] - want[ok]) / want[ok]
    print(f"start residual model {model}: {ok.sum()} windows, max rel err {rel.max():.3g}")
    assert rel.max() <= 1e-10
