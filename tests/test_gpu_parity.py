"""GPU parity: the HIP kernels (through the C ABI) against the FP64 oracle and against
the host build of the same lane logic.  The bar (tests/parity.py): accept/reject status
agreement >= 99.9 %; on fits both sides accept and converge, x, y within 1e-3 px and every
other parameter and record field (deviations, chi, integral, integral deviation, sd1, sd2
and their deviations, theta) within 1e-4 for >= 99.9 % of the fits; every both-accepted fit
(including those that hit the iteration cap or GSL_ENOPROG, where the reference itself is
not reproducible under a one-ulp change of its start) at the same bar for the separable
models and 99.8 % for the ellipsoidal model.  A fit's result must not depend on its batch:
the straggler kernel's groups compute the one-lane pass bit for bit, so results are
bit-identical with any hand-off threshold, none included."""
import sys

import numpy as np
import pytest

from localizer_amd import abi, native, synth
from oracle import pyoracle as po
from tests.emu import pyemu
from tests.parity import assert_parity, compare, format_stats, oracle_sigma_ulp

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def plain_cpu():
    return synth.generate(synth.PLAIN_CPU_512)


@pytest.fixture(scope="module")
def plain_gpu_sample():
    return synth.generate(synth.PLAIN_GPU_256.scaled(200))


def _check_sym(st):
    assert_parity(st, abi.MODEL_SYMMETRIC)


def _bytes_differ(a, b):
    """Per-fit flags: records that differ in any byte."""
    return np.any(a.view(np.uint8).reshape(a.size, -1) != b.view(np.uint8).reshape(b.size, -1), axis=1)


def test_gpu_symmetric_7x7_matches_oracle(plain_cpu):
    _say("case: symmetric 7x7, 100 frames 512x512, 2000 fits")
    frames, _, cands = plain_cpu
    with native.Context(abi.MODEL_SYMMETRIC, 1.3, default_half_width=3) as ctx:
        got = ctx.fit(frames, cands)
    ref = po.fit_batch(frames, cands, abi.MODEL_SYMMETRIC, 1.3, default_half_width=3)
    st = compare(ref, got, abi.MODEL_SYMMETRIC)
    _say(format_stats("gpu vs oracle 7x7", st))
    _check_sym(st)
    assert st["xy_frac_within_1e-3"] == 1.0


def test_gpu_symmetric_9x9_matches_oracle(plain_gpu_sample):
    _say("case: symmetric 9x9, 200 frames 256x256, 10000 fits")
    frames, _, cands = plain_gpu_sample
    with native.Context(abi.MODEL_SYMMETRIC, 1.35, default_half_width=4) as ctx:
        got = ctx.fit(frames, cands)
    ref = po.fit_batch(frames, cands, abi.MODEL_SYMMETRIC, 1.35, default_half_width=4)
    st = compare(ref, got, abi.MODEL_SYMMETRIC)
    _say(format_stats("gpu vs oracle 9x9", st))
    _check_sym(st)


def test_gpu_matches_host_build_of_lane_logic(plain_gpu_sample):
    _say("case: gpu vs host emulation, 9x9")
    frames, _, cands = plain_gpu_sample
    with native.Context(abi.MODEL_SYMMETRIC, 1.35, default_half_width=4) as ctx:
        got = ctx.fit(frames, cands)
    emu = pyemu.fit_batch(frames, cands, abi.MODEL_SYMMETRIC, 1.35, default_half_width=4)
    st = compare(emu, got, abi.MODEL_SYMMETRIC)
    _say(format_stats("gpu vs emulation 9x9", st))
    assert_parity(st, abi.MODEL_SYMMETRIC)
    assert st["iterations_equal"] >= 0.999


def test_gpu_symmetric_psf_filter(plain_gpu_sample):
    _say("case: ellipsoidal + symmetric-PSF filter, 10000 fits")
    frames, _, cands = plain_gpu_sample
    with native.Context(abi.MODEL_ELLIPSOIDAL_SYMMETRIC_PSF, 1.35, default_half_width=4) as ctx:
        got = ctx.fit(frames, cands)
    ref = po.fit_batch(frames, cands, abi.MODEL_ELLIPSOIDAL_SYMMETRIC_PSF, 1.35, default_half_width=4)
    st = compare(ref, got, abi.MODEL_ELLIPSOIDAL_SYMMETRIC_PSF)
    _say(format_stats("gpu vs oracle symmetric-PSF filter", st))
    spread = compare(ref, oracle_sigma_ulp(po, frames, cands, abi.MODEL_ELLIPSOIDAL_SYMMETRIC_PSF, 1.35,
                                           default_half_width=4), abi.MODEL_ELLIPSOIDAL_SYMMETRIC_PSF)
    assert_parity(st, abi.MODEL_ELLIPSOIDAL_SYMMETRIC_PSF, spread=spread)
    assert set(np.unique(got["status"])) <= {0, 2, 3, 4, 5, 6, 7}
    assert {abi.REJECT_ANISOTROPIC, abi.REJECT_PSF_WIDTH} & set(np.unique(got["status"]))


def test_gpu_fixed_width_matches_oracle(plain_gpu_sample):
    _say("case: fixed width 9x9, 200 frames 256x256, 10000 fits")
    frames, _, cands = plain_gpu_sample
    with native.Context(abi.MODEL_FIXED_WIDTH, 1.35, default_half_width=4) as ctx:
        got = ctx.fit(frames, cands)
    ref = po.fit_batch(frames, cands, abi.MODEL_FIXED_WIDTH, 1.35, default_half_width=4)
    st = compare(ref, got, abi.MODEL_FIXED_WIDTH)
    _say(format_stats("gpu vs oracle fixed width", st))
    assert_parity(st, abi.MODEL_FIXED_WIDTH)
    assert set(np.unique(got["status"])) <= {0, 2, 3, 9}


@pytest.fixture(scope="module")
def plain_100k():
    w = synth.PLAIN_GPU_256.scaled(2000)
    frames, _, cands = synth.generate(w, nthreads=16)
    ref = po.fit_batch(frames, cands, abi.MODEL_SYMMETRIC, 1.35, default_half_width=4)
    with native.Context(abi.MODEL_SYMMETRIC, 1.35, default_half_width=4) as ctx:
        ctx.set_option(abi.OPT_STRAGGLER_TRIALS, 0)
        unsplit = ctx.fit(frames, cands)
    return frames, cands, ref, unsplit


@pytest.mark.parametrize("trials", [8, 24])
def test_gpu_straggler_groups_11x11(trials):
    """11 x 11 windows (half-width 5) run the straggler build for windows up to 11 x 11; its
    groups give the records of the launch without hand-off bit for bit."""
    w = synth.LEARNED_INIT_11
    rois, _, cands = synth.generate_rois(w, 20000, 5, 1.0)
    _say(f"case: symmetric 11x11, 20k fits, hand-off after {trials} trials")
    with native.Context(abi.MODEL_SYMMETRIC, w.initial_psf_width, default_half_width=5) as ctx:
        ctx.set_option(abi.OPT_STRAGGLER_TRIALS, 0)
        unsplit = ctx.fit(rois, cands)
        ctx.set_option(abi.OPT_STRAGGLER_TRIALS, trials)
        got = ctx.fit(rois, cands)
    handed = int(np.sum(unsplit["trials"] > trials))
    diff = _bytes_differ(got, unsplit)
    _say(f"{handed} fits handed off; records differing from the launch without hand-off: {int(diff.sum())}")
    assert handed > 100
    assert not diff.any()


@pytest.mark.parametrize("rounds", [2, 3])
def test_gpu_parking_rounds(plain_100k, rounds):
    """LZR_OPT_ROUNDS > 1 launches the persistent kernel's build with the parking and resume
    paths: a drained wave with few busy lanes parks their fits, and the next launch continues
    them.  One round (the default) launches the build without those paths.  Both give the
    records of the launch without hand-off bit for bit."""
    frames, cands, ref, unsplit = plain_100k
    _say(f"case: symmetric 9x9, 100k fits, {rounds} rounds")
    with native.Context(abi.MODEL_SYMMETRIC, 1.35, default_half_width=4) as ctx:
        ctx.set_option(abi.OPT_ROUNDS, rounds)
        got = ctx.fit(frames, cands)
        assert ctx.get_option(abi.OPT_ROUNDS) == rounds
    diff = _bytes_differ(got, unsplit)
    _say(f"records differing from the one-round launch without hand-off: {int(diff.sum())}")
    assert not diff.any()


@pytest.mark.parametrize("nframes", [1, 16, 20, 21, 300])
def test_gpu_small_calls_hand_off_at_once(plain_100k, nframes):
    """A call of at most 1024 candidates (frames of 50: up to 20 frames) hands every fit to
    the straggler groups after its first trial unless the threshold was set; 21 frames (1050
    candidates) keep the default threshold.  Every call launches only as many waves as it has
    fits.  Every size gives the records of the launch without hand-off bit for bit."""
    frames, cands, ref, unsplit = plain_100k
    sel = cands["frame"] < nframes
    c = cands[sel]
    _say(f"case: symmetric 9x9, {nframes} frames, {c.size} fits, default hand-off policy")
    with native.Context(abi.MODEL_SYMMETRIC, 1.35, default_half_width=4) as ctx:
        got = ctx.fit(frames[:nframes], c)
        assert ctx.get_option(abi.OPT_STRAGGLER_TRIALS) == 48
    diff = _bytes_differ(got, unsplit[sel])
    _say(f"records differing from the launch without hand-off: {int(diff.sum())}")
    assert not diff.any()


@pytest.mark.parametrize("trials", [4, 11, 17, 24, 48])
def test_gpu_straggler_group_sizes(plain_100k, trials):
    """Fits still running after `trials` trials go to the straggler kernel, whose group size
    follows their count (at most two turns of the launch's 65536 lanes on a 256-CU MI355X):
    after 4 / 11 / 17 / 24 / 48 trials about 99k / 46k / 25k / 13k / 1.7k of the 100k fits
    are handed off, which selects one lane per fit (the two-wave build; this model skips
    2-lane groups) for the first two and groups of 4, 8 and 16 lanes for the others.
    Every group size must give the fits of a launch without hand-off bit for bit, and match
    the oracle."""
    frames, cands, ref, unsplit = plain_100k
    _say(f"case: symmetric 9x9, 100k fits, hand-off after {trials} trials")
    with native.Context(abi.MODEL_SYMMETRIC, 1.35, default_half_width=4) as ctx:
        ctx.set_option(abi.OPT_STRAGGLER_TRIALS, trials)
        got = ctx.fit(frames, cands)
        again = ctx.fit(frames, cands)
    handed = int(np.sum(ref["trials"] > trials))
    st = compare(ref, got, abi.MODEL_SYMMETRIC)
    _say(format_stats(f"gpu vs oracle, ~{handed} fits handed off", st))
    diff = _bytes_differ(got, unsplit)
    _say(f"records differing from the launch without hand-off: {int(diff.sum())}")
    assert not diff.any()
    assert np.array_equal(got.view(np.uint8), again.view(np.uint8))
    _check_sym(st)


def test_gpu_mixed_models_and_windows():
    _say("case: mixed models (symmetric, ellipsoidal, fixed width) and windows h=2..7, 20k fits")
    w = synth.MIXED_2M.scaled(400)
    frames, _, cands = synth.generate(w)
    cands["model"][::7] = abi.MODEL_FIXED_WIDTH
    with native.Context(abi.MODEL_SYMMETRIC, 1.5) as ctx:
        ctx.set_option(abi.OPT_MAX_HALF_WIDTH, 7)
        got = ctx.fit(frames, cands)
    ref = po.fit_batch(frames, cands, abi.MODEL_SYMMETRIC, 1.5)
    other = oracle_sigma_ulp(po, frames, cands, abi.MODEL_SYMMETRIC, 1.5)
    assert np.all(got["model"] == ref["model"])
    for m in (abi.MODEL_SYMMETRIC, abi.MODEL_ELLIPSOIDAL, abi.MODEL_FIXED_WIDTH):
        sel = cands["model"] == m
        st = compare(ref[sel], got[sel], m)
        _say(format_stats(f"gpu vs oracle mixed model {m}", st))
        assert_parity(st, m, spread=compare(ref[sel], other[sel], m))


@pytest.fixture(scope="module")
def ellipsoidal_100k(plain_100k):
    frames, cands, _, _ = plain_100k
    ref = po.fit_batch(frames, cands, abi.MODEL_ELLIPSOIDAL, 1.35, default_half_width=4)
    spread = compare(ref, oracle_sigma_ulp(po, frames, cands, abi.MODEL_ELLIPSOIDAL, 1.35, default_half_width=4),
                     abi.MODEL_ELLIPSOIDAL)
    with native.Context(abi.MODEL_ELLIPSOIDAL, 1.35, default_half_width=4) as ctx:
        ctx.set_option(abi.OPT_STRAGGLER_TRIALS, 0)
        unsplit = ctx.fit(frames, cands)
    return frames, cands, ref, unsplit, spread


def test_gpu_ellipsoidal_matches_oracle(ellipsoidal_100k):
    """The ellipsoidal model on 100k 9x9 fits (default hand-off) to the parity bar, with the
    reference's own reproducibility measured on the same fits."""
    frames, cands, ref, unsplit, spread = ellipsoidal_100k
    _say("case: ellipsoidal 9x9, 100k fits")
    with native.Context(abi.MODEL_ELLIPSOIDAL, 1.35, default_half_width=4) as ctx:
        got = ctx.fit(frames, cands)
    st = compare(ref, got, abi.MODEL_ELLIPSOIDAL)
    _say(format_stats("gpu vs oracle ellipsoidal", st))
    assert_parity(st, abi.MODEL_ELLIPSOIDAL, spread=spread)
    assert not _bytes_differ(got, unsplit).any()


@pytest.mark.parametrize("trials", [3, 8, 14, 23, 48])
def test_gpu_ellipsoidal_straggler_group_sizes(ellipsoidal_100k, trials):
    """The ellipsoidal straggler pass (rows of the window split over the group, their sums
    folded in row order) at every group size: after 3 / 8 / 14 / 23 / 48 trials about 100k /
    42k / 25k / 13k / 3k of the 100k fits are handed off (groups of 1, 2, 4, 8, 16 lanes on
    a 256-CU MI355X).  Bit-identical to the launch without hand-off, and against the oracle as
    the ellipsoidal parity test."""
    frames, cands, ref, unsplit, spread = ellipsoidal_100k
    _say(f"case: ellipsoidal 9x9, 100k fits, hand-off after {trials} trials")
    with native.Context(abi.MODEL_ELLIPSOIDAL, 1.35, default_half_width=4) as ctx:
        ctx.set_option(abi.OPT_STRAGGLER_TRIALS, trials)
        got = ctx.fit(frames, cands)
        again = ctx.fit(frames, cands)
    st = compare(ref, got, abi.MODEL_ELLIPSOIDAL)
    _say(format_stats(f"gpu vs oracle ellipsoidal, ~{int(np.sum(ref['trials'] > trials))} fits handed off", st))
    diff = _bytes_differ(got, unsplit)
    _say(f"records differing from the launch without hand-off: {int(diff.sum())}")
    assert not diff.any()
    assert np.array_equal(got.view(np.uint8), again.view(np.uint8))
    assert_parity(st, abi.MODEL_ELLIPSOIDAL, spread=spread)


def test_gpu_mixed_windows_bit_identical():
    """Lanes with different window sizes share a wave (its loop runs to the largest window),
    and which fits share a wave depends on timing: repeated calls must be bit-identical for
    every model.  (A product 0 * inf outside a lane's window once made ~0.1 % of the
    ellipsoidal fits depend on their wave neighbours.)"""
    _say("case: mixed windows h=2..7, three models, 80k fits, three calls")
    w = synth.MIXED_2M.scaled(1600)
    frames, _, cands = synth.generate(w, nthreads=16)
    cands["model"][::7] = abi.MODEL_FIXED_WIDTH
    outs = []
    for _ in range(3):
        with native.Context(abi.MODEL_SYMMETRIC, 1.5) as ctx:
            ctx.set_option(abi.OPT_MAX_HALF_WIDTH, 7)
            outs.append(ctx.fit(frames, cands))
    for o in outs[1:]:
        diff = np.any(o.view(np.uint8).reshape(o.size, -1) != outs[0].view(np.uint8).reshape(o.size, -1), axis=1)
        assert not diff.any(), (int(diff.sum()), np.unique(cands["model"][diff]))


def test_gpu_edges_invalid_and_empty():
    _say("case: edge / invalid / empty")
    frames = np.full((2, 32, 40), 100, np.uint16)
    frames[0, 10:13, 10:13] = 400
    c = np.zeros(6, abi.CANDIDATE_DTYPE)
    c["x"] = [11, 2.0, 37.5, 11, 11, 11]
    c["y"] = [11, 11, 11, 1.0, 11, 11]
    c["intensity"] = 300
    c["background"] = 100
    c["frame"] = [0, 0, 0, 0, 5, 0]
    c["half_width"] = [3, 3, 3, 3, 3, 16]
    c["model"] = -1
    with native.Context(abi.MODEL_SYMMETRIC, 1.3, default_half_width=3) as ctx:
        got = ctx.fit(frames, c)
        empty = ctx.fit(frames, c[:0])
    ref = po.fit_batch(frames, c[:4], abi.MODEL_SYMMETRIC, 1.3, default_half_width=3)
    assert list(got["status"][:4]) == list(ref["status"])
    assert list(got["status"][1:4]) == [abi.REJECT_EDGE] * 3
    assert got["status"][4] == abi.REJECT_INVALID and got["status"][5] == abi.REJECT_INVALID
    assert empty.size == 0


def test_gpu_device_entry_matches_host_entry(plain_cpu):
    _say("case: device-resident entry point")
    import torch

    frames, _, cands = plain_cpu
    d_frames = torch.from_numpy(frames.view(np.int16)).cuda()
    d_cands = torch.from_numpy(cands.view(np.uint8)).cuda()
    d_out = torch.zeros(cands.size * abi.FIT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    with native.Context(abi.MODEL_SYMMETRIC, 1.3, default_half_width=3) as ctx:
        host = ctx.fit(frames, cands)
        stream = torch.cuda.current_stream().cuda_stream
        for _ in range(2):  # repeat: the queue state is reset per call
            ctx.fit_device(d_frames.data_ptr(), 512, 512, frames.shape[0], d_cands.data_ptr(), cands.size,
                           d_out.data_ptr(), stream)
        torch.cuda.synchronize()
    dev = d_out.cpu().numpy().view(abi.FIT_DTYPE)
    assert np.array_equal(dev["status"], host["status"])
    assert np.array_equal(dev["params"], host["params"])


def test_gpu_results_independent_of_batching():
    """The mixed workload (windows 5x5 to 15x15, three models) fitted as one call and as four
    blocks of frames: every record bit-identical (a fit's result depends on its own spot only,
    source/Localization.cpp:107-212 fits each particle on its own)."""
    _say("case: mixed windows and models, 80k fits, one call vs four blocks")
    w = synth.MIXED_2M.scaled(1600)
    frames, _, cands = synth.generate(w, nthreads=16)
    cands["model"][::7] = abi.MODEL_FIXED_WIDTH
    with native.Context(abi.MODEL_SYMMETRIC, 1.5) as ctx:
        ctx.set_option(abi.OPT_MAX_HALF_WIDTH, 7)
        whole = ctx.fit(frames, cands)
        parts = []
        edges = np.linspace(0, cands.size, 5).astype(int)
        for a, b in zip(edges[:-1], edges[1:]):
            parts.append(ctx.fit(frames, cands[a:b]))
    blocks = np.concatenate(parts)
    diff = _bytes_differ(whole, blocks)
    _say(f"records differing between one call and four blocks: {int(diff.sum())}")
    assert not diff.any()


@pytest.fixture(scope="module")
def mixed_100k():
    w = synth.MIXED_2M.scaled(2000)
    frames, _, cands = synth.generate(w, nthreads=16)
    ref = po.fit_batch(frames, cands, abi.MODEL_SYMMETRIC, 1.5)
    other = oracle_sigma_ulp(po, frames, cands, abi.MODEL_SYMMETRIC, 1.5)
    return frames, cands, ref, other


def test_gpu_mixed_100k_matches_oracle(mixed_100k):
    """The batched uneven workload (h ~ U{2..7}, 60 % symmetric / 40 % ellipsoidal) at 100k
    ROIs in one call: each model class to the parity bar."""
    frames, cands, ref, other = mixed_100k
    _say(f"case: mixed 100k, {cands.size} fits")
    with native.Context(abi.MODEL_SYMMETRIC, 1.5) as ctx:
        ctx.set_option(abi.OPT_MAX_HALF_WIDTH, 7)
        got = ctx.fit(frames, cands)
    for m in (abi.MODEL_SYMMETRIC, abi.MODEL_ELLIPSOIDAL):
        sel = got["model"] == m
        st = compare(ref[sel], got[sel], m)
        _say(format_stats(f"gpu vs oracle mixed 100k model {m}", st))
        assert_parity(st, m, spread=compare(ref[sel], other[sel], m))


def test_gpu_float_frames(mixed_100k):
    """The float entry points (lzr_fit_batch_f32): integer frames passed as float give the
    uint16 records bit for bit (mixed models, windows 5x5..15x15, straggler hand-offs
    included); non-integer frames (1.375 v + 0.25, exact in FP32) meet the bar against the
    FP64 oracle fitting the same FP64 frames."""
    frames, cands, _, _ = mixed_100k
    frames, cands = frames[:400], cands[cands["frame"] < 400]
    _say(f"case: float frames, {cands.size} mixed fits")
    with native.Context(abi.MODEL_SYMMETRIC, 1.5) as ctx:
        ctx.set_option(abi.OPT_MAX_HALF_WIDTH, 7)
        u16 = ctx.fit(frames, cands)
        f32 = ctx.fit(frames.astype(np.float32), cands)
        assert u16.tobytes() == f32.tobytes()
        # the same spots on an affine pixel scale, with the candidates a finder would report
        ff = frames.astype(np.float64) * 1.375 + 0.25
        cs = cands.copy()
        cs["intensity"] *= 1.375
        cs["background"] = cs["background"] * 1.375 + 0.25
        got = ctx.fit(ff.astype(np.float32), cs)
    ref = po.fit_batch(ff, cs, abi.MODEL_SYMMETRIC, 1.5)
    other = oracle_sigma_ulp(po, ff, cs, abi.MODEL_SYMMETRIC, 1.5)
    for m in (abi.MODEL_SYMMETRIC, abi.MODEL_ELLIPSOIDAL):
        sel = got["model"] == m
        st = compare(ref[sel], got[sel], m)
        _say(format_stats(f"gpu float frames vs FP64 oracle, model {m}", st))
        assert_parity(st, m, spread=compare(ref[sel], other[sel], m), slack=1)
