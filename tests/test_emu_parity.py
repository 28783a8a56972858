"""The GPU lane logic (localizer_amd/csrc/lzr_lane.h), compiled for the host, against the
FP64 oracle.  This checks the normal-equation restatement of lmsder and its FP32/FP64
split on the CPU; the same code runs on the GPU in tests/test_gpu_parity.py."""
import numpy as np
import pytest

from localizer_amd import abi, synth
from oracle import pyoracle as po
from tests.emu import pyemu
from tests.parity import compare, format_stats


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
    assert st["status_agreement"] >= 0.999
    assert st["xy_frac_within_1e-3"] == 1.0
    for k in ("A", "r", "b"):
        assert st[f"{k}_frac_within_1e-4"] >= 0.995


def test_emulated_lanes_match_oracle_9x9(plain_gpu_sample):
    frames, _, cands = plain_gpu_sample
    ref = po.fit_batch(frames, cands, abi.MODEL_SYMMETRIC, 1.35, default_half_width=4)
    got = pyemu.fit_batch(frames, cands, abi.MODEL_SYMMETRIC, 1.35, default_half_width=4)
    st = compare(ref, got, abi.MODEL_SYMMETRIC)
    print(format_stats("emu 9x9", st))
    assert st["status_agreement"] >= 0.999
    assert st["xy_frac_within_1e-3"] >= 0.999
    for k in ("A", "r", "b"):
        assert st[f"{k}_frac_within_1e-4"] >= 0.995


def test_emulated_lanes_match_oracle_ellipsoidal(plain_gpu_sample):
    frames, _, cands = plain_gpu_sample
    ref = po.fit_batch(frames, cands, abi.MODEL_ELLIPSOIDAL, 1.35, default_half_width=4)
    got = pyemu.fit_batch(frames, cands, abi.MODEL_ELLIPSOIDAL, 1.35, default_half_width=4)
    st = compare(ref, got, abi.MODEL_ELLIPSOIDAL)
    print(format_stats("emu ellipsoidal 9x9", st))
    assert st["status_agreement"] >= 0.995
    # fits that converge in both (test_delta, same iteration count) agree tightly
    ok = (ref["lm_status"] == 0) & (got["lm_status"] == 0) & (ref["status"] == 0) & (got["status"] == 0)
    dxy = np.abs(ref["params"][ok][:, 3:5] - got["params"][ok][:, 3:5]).max(axis=1)
    assert np.mean(dxy <= 1e-3) >= 0.995


def test_emulated_lanes_match_oracle_fixed_width(plain_gpu_sample):
    frames, _, cands = plain_gpu_sample
    ref = po.fit_batch(frames, cands, abi.MODEL_FIXED_WIDTH, 1.35, default_half_width=4)
    got = pyemu.fit_batch(frames, cands, abi.MODEL_FIXED_WIDTH, 1.35, default_half_width=4)
    st = compare(ref, got, abi.MODEL_FIXED_WIDTH)
    print(format_stats("emu fixed width 9x9", st))
    assert st["status_agreement"] >= 0.999
    assert st["xy_frac_within_1e-3"] >= 0.999
    for k in ("A", "b"):
        assert st[f"{k}_frac_within_1e-4"] >= 0.995
    acc = (ref["status"] == 0) & (got["status"] == 0)
    np.testing.assert_allclose(got["integral"][acc], ref["integral"][acc], rtol=1e-3)
    np.testing.assert_allclose(got["deviations"][acc][:, :4], ref["deviations"][acc][:, :4], rtol=1e-2)
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
