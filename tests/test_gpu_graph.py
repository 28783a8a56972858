"""Frame-block graphs and the autotuner (GPU).  A captured fit replayed on refilled
buffers must give the eager call's results for every block; the autotuner's choice
must not change results."""
import sys

import numpy as np
import pytest

from localizer_amd import abi, device, native, synth

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_gpu_graph_replays_match_eager_calls():
    import torch

    _say("case: frame-block graph, mixed models and windows, 4 blocks of uneven size")
    w = synth.MIXED_2M.scaled(48)
    frames, _, cands = synth.generate(w)
    rng = np.random.default_rng(5)
    blocks = []
    edges = [0, 10, 22, 31, 48]
    for b in range(4):
        f0, f1 = edges[b], edges[b + 1]
        sel = (cands["frame"] >= f0) & (cands["frame"] < f1)
        c = cands[sel].copy()
        c["frame"] -= f0
        c = c[rng.permutation(len(c))][: len(c) - b * 37]  # uneven counts
        blocks.append((frames[f0:f1], c))
    max_f = max(len(f) for f, _ in blocks)
    max_c = max(len(c) for _, c in blocks)
    st_frames = device.empty(np.uint16, max_f * w.ny * w.nx)
    st_cands = device.empty(abi.CANDIDATE_DTYPE, max_c)
    st_count = torch.zeros(1, dtype=torch.int32, device="cuda")
    st_out = device.empty(abi.FIT_DTYPE, max_c)
    s = device.stream_handle()
    with native.Context(abi.MODEL_SYMMETRIC, w.initial_psf_width) as ctx:
        ctx.set_option(abi.OPT_MAX_HALF_WIDTH, 7)
        g = ctx.graph(st_frames.ptr, w.nx, w.ny, max_f, st_cands.ptr, st_count.data_ptr(), max_c, st_out.ptr)
        for rep in range(2):
            for b, (f, c) in enumerate(blocks):
                eager = ctx.fit(f, c)  # eager work between replays
                st_frames.t[: f.size * 2].copy_(torch.from_numpy(f.reshape(-1).view(np.uint8)))
                st_cands.t[: c.size * abi.CANDIDATE_DTYPE.itemsize].copy_(torch.from_numpy(c.view(np.uint8)))
                st_count.fill_(len(c))
                torch.cuda.synchronize()
                g.launch(s)
                torch.cuda.synchronize()
                got = st_out.t.cpu().numpy().view(abi.FIT_DTYPE)[: len(c)]
                assert np.array_equal(got["status"], eager["status"]), (rep, b)
                assert np.array_equal(got["params"], eager["params"]), (rep, b)
                assert np.array_equal(got["iterations"], eager["iterations"]), (rep, b)
        _say(f"graph: {2 * len(blocks)} replays identical to eager calls")
        g.close()


def test_gpu_graph_survives_larger_eager_calls():
    """A graph owns its work buffers: an eager call on the same context with twice the graph's
    capacity (which grows the context's own queue and hand-off buffers) must not move what the
    graph's nodes point to.  Graph of N candidates, eager call of 2N, replay, compare."""
    import torch

    _say("case: graph of N, eager 2N in between, replay")
    w = synth.PLAIN_GPU_256.scaled(80)
    frames, _, cands = synth.generate(w)
    n = cands.size // 2
    part = cands[:n]
    st_frames = device.empty(np.uint16, frames.size)
    st_cands = device.empty(abi.CANDIDATE_DTYPE, n)
    st_count = torch.zeros(1, dtype=torch.int32, device="cuda")
    st_out = device.empty(abi.FIT_DTYPE, n)
    s = device.stream_handle()
    with native.Context(abi.MODEL_SYMMETRIC, w.initial_psf_width, default_half_width=4) as ctx:
        ctx.set_option(abi.OPT_STRAGGLER_TRIALS, 11)  # many hand-offs: the hand-off region is in use
        g = ctx.graph(st_frames.ptr, w.nx, w.ny, frames.shape[0], st_cands.ptr, st_count.data_ptr(), n, st_out.ptr)
        big = ctx.fit(frames, cands)  # 2N: the context's buffers are reallocated
        small = ctx.fit(frames, part)
        st_frames.t.copy_(torch.from_numpy(frames.reshape(-1).view(np.uint8)))
        st_cands.t.copy_(torch.from_numpy(part.view(np.uint8)))
        st_count.fill_(n)
        torch.cuda.synchronize()
        for _ in range(2):
            g.launch(s)
            torch.cuda.synchronize()
            got = st_out.t.cpu().numpy().view(abi.FIT_DTYPE)[:n]
            assert np.array_equal(got.view(np.uint8), small.view(np.uint8))
            ctx.fit(frames, cands)  # and again after another large eager call
        assert np.array_equal(big[:n].view(np.uint8), small.view(np.uint8))  # batch independence
        g.close()


def test_gpu_autotuner_keeps_results():
    _say("case: autotuner on a mixed batch")
    w = synth.MIXED_2M.scaled(60)
    frames, _, cands = synth.generate(w)
    d_frames, d_cands = device.upload(frames), device.upload(cands)
    d_out = device.empty(abi.FIT_DTYPE, cands.size)
    s = device.stream_handle()
    with native.Context(abi.MODEL_SYMMETRIC, w.initial_psf_width) as ctx:
        ctx.set_option(abi.OPT_MAX_HALF_WIDTH, 7)
        ref = ctx.fit(frames, cands)
        best = ctx.autotune_device(d_frames.ptr, w.nx, w.ny, w.nframes, d_cands.ptr, cands.size, d_out.ptr, s)
        chosen = {k: ctx.get_option(v) for k, v in (("occupancy", abi.OPT_OCCUPANCY),
                                                       ("waves_per_cu", abi.OPT_WAVES_PER_CU),
                                                       ("rounds", abi.OPT_ROUNDS))}
        _say(f"autotune: best {best:.3f} ms with {chosen}")
        # the context's own (automatic, 0) configuration is a candidate too
        assert chosen["occupancy"] in (0, 1, 2) and chosen["rounds"] in (1, 2)
        got = ctx.fit(frames, cands)
    assert np.array_equal(got["status"], ref["status"])
    np.testing.assert_allclose(got["params"], ref["params"], rtol=1e-12, atol=0)
