"""The allocation-free entry point (lzr_fit_batch_device_noalloc) against lzr_fit_batch_device
through the C ABI: two calls in a row, the second smaller, each bit-identical to the allocating
entry's result.  With one trial before hand-off nearly every fit of the first call is handed
off, more than the fixed array's hand-off entries hold: the fits that find them taken finish in
the first kernel, with the same results."""
import ctypes
import sys

import numpy as np
import pytest

from localizer_amd import abi, device, native, synth

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _run(entry, ctx, frames, cands, stream):
    import torch

    d_frames = torch.from_numpy(frames.reshape(-1).view(np.int16)).cuda()
    d_cands = torch.from_numpy(cands.view(np.uint8)).cuda()
    out = device.empty(abi.FIT_DTYPE, cands.size)
    nf, ny, nx = frames.shape
    rc = entry(ctx._h, ctypes.c_void_p(d_frames.data_ptr()), nx, ny, nf, ctypes.c_void_p(d_cands.data_ptr()),
               cands.size, ctypes.c_void_p(out.ptr), ctypes.c_void_p(stream))
    assert rc == 0, native._error(ctx._h)
    torch.cuda.synchronize()
    return out.t.cpu().numpy().view(abi.FIT_DTYPE)[: cands.size].copy()


@pytest.mark.parametrize("trials", [48, 1])
def test_noalloc_matches_allocating_entry(trials):
    L = native.lib()
    s = device.stream_handle()
    w = synth.MIXED_2M.scaled(4000)  # 200k fits, windows 5x5 to 15x15, three models
    frames, _, cands = synth.generate(w, nthreads=16)
    cands["model"][::7] = abi.MODEL_FIXED_WIDTH
    small = cands[cands["frame"] < 1000]
    _say(f"case: no-alloc entry, hand-off after {trials} trials, {cands.size} then {small.size} fits")
    with native.Context(abi.MODEL_SYMMETRIC, 1.5) as ctx:
        ctx.set_option(abi.OPT_MAX_HALF_WIDTH, 7)
        ctx.set_option(abi.OPT_STRAGGLER_TRIALS, trials)
        for c in (cands, small):
            want = _run(L.lzr_fit_batch_device, ctx, frames, c, s)
            got = _run(L.lzr_fit_batch_device_noalloc, ctx, frames, c, s)
            diff = np.any(got.view(np.uint8).reshape(c.size, -1) != want.view(np.uint8).reshape(c.size, -1), axis=1)
            _say(f"  {c.size} fits: {int(diff.sum())} records differ, accepted {int((got['status'] == 0).sum())}")
            assert not diff.any()


def test_noalloc_rejects_batches_beyond_its_maximum():
    L = native.lib()
    with native.Context(abi.MODEL_SYMMETRIC, 1.35, default_half_width=4) as ctx:
        rc = L.lzr_fit_batch_device_noalloc(ctx._h, ctypes.c_void_p(1), 16, 16, 1, ctypes.c_void_p(1),
                                            (2 << 20) + 1, ctypes.c_void_p(1), None)
        assert rc == -1
        assert "exceed" in native._error(ctx._h)
