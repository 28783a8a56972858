"""GPU tests of the host batcher (lzr_batcher_*): the reference controller's call pattern --
worker threads sharing one fitter, one fit_positions call per frame
(source/PALMAnalysis.cpp:204-252, call at :234) -- gathered into device batches.  Every
record must equal the record a one-frame lzr_fit_batch call returns, bit for bit."""
import threading

import numpy as np
import pytest

from localizer_amd import abi, native, synth

pytestmark = pytest.mark.gpu

W = synth.PLAIN_GPU_256


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def frames_and_lists():
    frames, _, cands = synth.generate(W.scaled(160))
    lists = []
    rng = np.random.default_rng(3)
    for f in range(frames.shape[0]):
        c = cands[cands["frame"] == f].copy()
        c["frame"] = 0
        if f % 7 == 0 and c.size > 3:
            c["frame"][0] = 3          # a request for another frame: invalid in a one-frame call
            c["x"][1] = 1.0            # window outside the frame: edge rejection
            c["half_width"][2] = 40    # beyond the largest window: invalid
        if f % 11 == 0:
            c = c[:0]                  # a frame without candidates
        lists.append(c[rng.permutation(c.size)])
    return frames, lists


def _per_frame(ctx, frames, lists):
    return [ctx.fit(frames[f], lists[f]) if lists[f].size else np.zeros(0, abi.FIT_DTYPE) for f in range(len(lists))]


def _run_threads(batcher, frames, lists, nthreads, as_f64=False):
    out = [None] * len(lists)
    errors = []

    def work(t):
        try:
            for f in range(t, len(lists), nthreads):
                img = frames[f].astype(np.float64) if as_f64 else frames[f]
                out[f] = batcher.fit_frame(img, lists[f])
        except Exception as e:  # noqa: BLE001 -- reported below
            errors.append(repr(e))

    th = [threading.Thread(target=work, args=(t,)) for t in range(nthreads)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    return out


@pytest.mark.parametrize("nthreads,as_f64", [(16, False), (16, True), (1, False)])
def test_gpu_batcher_records_equal_one_frame_calls(frames_and_lists, nthreads, as_f64):
    frames, lists = frames_and_lists
    _say(f"case: batcher, {len(lists)} frames of {W.nx}x{W.ny}, {nthreads} threads, f64={as_f64}")
    with native.Context(abi.MODEL_SYMMETRIC, W.initial_psf_width) as ctx:
        want = _per_frame(ctx, frames, lists)
        with native.Batcher(ctx, W.nx, W.ny, max_frames=32, max_rois=4096, max_wait_us=2000) as b:
            got = _run_threads(b, frames, lists, nthreads, as_f64)
            st = b.stats()
    n = sum(x.size for x in want)
    for f in range(len(lists)):
        assert got[f].tobytes() == want[f].tobytes(), f"frame {f}"
    st_all = np.concatenate(want)["status"]
    _say(f"{n} records identical; {st['batches']} batches for {st['frames']} frames; statuses "
         f"{dict(zip(*np.unique(st_all, return_counts=True)))}")
    assert {abi.REJECT_INVALID, abi.REJECT_EDGE} <= set(st_all.tolist())
    if nthreads > 1:
        assert st["batches"] < st["frames"]  # calls were gathered


def test_gpu_batcher_errors(frames_and_lists):
    frames, lists = frames_and_lists
    _say("case: batcher error paths")
    f = next(i for i in range(len(lists)) if lists[i].size > 5)
    c = lists[f]
    with native.Context(abi.MODEL_SYMMETRIC, W.initial_psf_width) as ctx:
        want = ctx.fit(frames[f], c)
        with native.Batcher(ctx, W.nx, W.ny, max_frames=4, max_rois=256, max_wait_us=100) as b:
            with pytest.raises(native.LocalizerError, match="frame size"):
                b.fit_frame(frames[f][:, :-1], c)
            img = frames[f].astype(np.float64)
            k = next(i for i in range(c.size) if want["status"][i] == 0)
            xi, yi = int(c["x"][k]), int(c["y"][k])
            bad = img.copy()
            bad[yi, xi] += 0.5  # a non-integer pixel inside a fitting window
            with pytest.raises(native.LocalizerError, match="integer counts"):
                b.fit_frame(bad, c)
            ok = img.copy()
            ok[0, 0] = 0.5 if all(int(v) > 12 for v in c["x"]) and all(int(v) > 12 for v in c["y"]) else ok[0, 0]
            assert b.fit_frame(ok, c).tobytes() == want.tobytes()  # pixels outside every window are not read
            big = np.concatenate([c] * (256 // c.size + 1))
            with pytest.raises(native.LocalizerError, match="capacity"):
                b.fit_frame(frames[f], big)
            assert b.fit_frame(frames[f], c).tobytes() == want.tobytes()  # still usable after errors
