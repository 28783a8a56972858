"""GPU tests of the learned initializer: the network's forward pass and gradients
(HIP, BF16 MFMA) against the CPU oracle (oracle/pynet.py), start points feeding the
LM, and what a network trained on the GPU gains over the plain start."""
import sys

import numpy as np
import pytest

from localizer_amd import abi, device, native, netinit, synth
from oracle import pynet

pytestmark = pytest.mark.gpu

W = synth.LEARNED_INIT_11


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def rois():
    return netinit.simulate(W, 3000, 91)


def _oracle_inputs(rois, cands):
    patches, xoff, yoff, valid = pynet.patches_for(rois, cands)
    x, med, scale = pynet.prep(patches)
    return x, (med, scale, xoff, yoff, valid)


def test_gpu_net_forward_matches_oracle(rois):
    _say("case: network forward, 3000 patches + 2 outside the frame")
    frames, _, cands = rois
    cands = np.concatenate([cands, cands[:2]])
    cands["x"][-2:] = [2.0, 5.0]
    cands["y"][-2:] = [5.0, 9.5]
    n = len(cands)
    with native.Net(W.initial_psf_width, seed=7) as net:
        params = net.get_params()
        d_frames, d_cands = device.upload(frames), device.upload(cands)
        d_out = device.empty(np.float32, n * 5)
        net.forward_device(d_frames.ptr, 11, 11, len(frames), d_cands.ptr, n, d_out.ptr, device.stream_handle())
        got = d_out.numpy().reshape(n, 5)
    assert np.all(np.isnan(got[-2:]))
    x, _ = _oracle_inputs(frames, cands[:-2])
    ref = pynet.forward(params, x)["t"]
    err = np.abs(got[:-2] - ref)
    scale = np.abs(ref).max()
    _say(f"forward: max |gpu - oracle| = {err.max():.3g}, output scale {scale:.3g}, "
         f"99.9th pct {np.quantile(err, 0.999):.3g}")
    assert err.max() <= 2e-2 * scale
    assert np.quantile(err, 0.99) <= 2e-3 * scale


def test_gpu_net_gradients_match_oracle(rois):
    _say("case: network gradients, 1000-patch batch labelled by the GPU LM")
    frames, _, cands = rois
    cands = cands[:1000]
    n = len(cands)
    d_frames, d_cands = device.upload(frames), device.upload(cands)
    with native.Context(abi.MODEL_SYMMETRIC, W.initial_psf_width, default_half_width=5) as ctx:
        d_fits = netinit.label(ctx, d_frames, len(frames), d_cands, n)
    fits = d_fits.numpy()
    with native.Net(W.initial_psf_width, seed=11) as net:
        params = net.get_params()
        loss, g = net.train_step_device(d_frames.ptr, 11, 11, len(frames), d_cands.ptr, d_fits.ptr, n, 0.0,
                                        apply=False, want_grads=True, stream=device.stream_handle())
        after = net.get_params()
    assert np.array_equal(after, params)  # apply=0 leaves the weights alone
    x, stats = _oracle_inputs(frames, cands)
    tgt, wt = pynet.targets_from_fits(fits, stats, W.initial_psf_width)
    loss_ref, g_ref = pynet.backward(params, pynet.forward(params, x), tgt, wt)
    _say(f"loss gpu {loss:.6g} oracle {loss_ref:.6g}; used {int(wt.sum())} of {n}")
    assert abs(loss - loss_ref) <= 2e-3 * loss_ref
    for k, (o, shape) in abi.NET_OFF.items():
        m = int(np.prod(shape))
        a, b = g[o:o + m].astype(np.float64), g_ref[o:o + m]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        cos = float(a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30))
        _say(f"grad {k}: rel err {rel:.3g}, cosine {cos:.7f}")
        assert rel < 2e-2 and cos > 0.9998, k


def test_gpu_start_points_drive_the_lm(rois):
    _say("case: LM start points (skip / reference start / explicit start)")
    frames, truth, cands = rois
    n = 600
    cands = cands[:n]
    starts = np.zeros(n, abi.START_DTYPE)
    starts["width"][:200] = 0.0   # reference start
    starts["width"][200:400] = -1.0  # skip
    starts["width"][400:] = W.initial_psf_width  # explicit start = the reference start's values
    starts["amplitude"][400:] = cands["intensity"][400:]
    starts["x"][400:] = cands["x"][400:]
    starts["y"][400:] = cands["y"][400:]
    starts["background"][400:] = cands["background"][400:]
    d_frames, d_cands, d_starts = device.upload(frames), device.upload(cands), device.upload(starts)
    sentinel = np.zeros(n, abi.FIT_DTYPE)
    sentinel["status"] = 77
    d_out = device.upload(sentinel)
    with native.Context(abi.MODEL_SYMMETRIC, W.initial_psf_width, default_half_width=5) as ctx:
        ctx.fit_device_start(d_frames.ptr, 11, 11, len(frames), d_cands.ptr, d_starts.ptr, n, d_out.ptr,
                             device.stream_handle())
        plain = ctx.fit(frames, cands)
    got = d_out.numpy()
    assert np.all(got["status"][200:400] == 77)  # skipped records untouched
    keep = np.r_[0:200, 400:n]
    assert np.array_equal(got["status"][keep], plain["status"][keep])
    assert np.array_equal(got["params"][keep], plain["params"][keep])


def test_gpu_trained_network_cuts_lm_work():
    """Train on the GPU, then compare net-start LM (residual fallback) against the plain
    LM on fresh ROIs: fewer trial evaluations, the same fits to 1e-3 px where both
    accept.  (The LM's convergence is linear -- the reference's width column is the
    true derivative / sqrt(2) -- so a better start saves a logarithmic number of steps.)"""
    _say("case: train 1048576 ROIs (512 steps x 2048), evaluate 20000 fresh ROIs")
    with native.Net(W.initial_psf_width, seed=3) as net:
        stats = netinit.train(net, W, 1 << 20, epochs=1.0, batch=2048, lr=2e-3, log=None)
        _say(f"train: {stats['steps']} steps in {stats['train_s']:.2f} s, losses {stats['losses']}")
        assert stats["losses"][-1][1] < 0.3 * stats["losses"][0][1]
        frames, _, cands = netinit.simulate(W, 20000, W.extra["eval_stream"])
        n = len(cands)
        d_frames, d_cands = device.upload(frames), device.upload(cands)
        d_plain = device.empty(abi.FIT_DTYPE, n)
        d_net = device.empty(abi.FIT_DTYPE, n)
        d_first = device.empty(abi.FIT_DTYPE, n)
        d_starts = device.empty(abi.START_DTYPE, n)
        d_work = device.empty(abi.START_DTYPE, n)
        s = device.stream_handle()
        with native.Context(abi.MODEL_SYMMETRIC, W.initial_psf_width, default_half_width=5) as ctx:
            ctx.fit_device(d_frames.ptr, 11, 11, n, d_cands.ptr, n, d_plain.ptr, s)
            net.predict_device(d_frames.ptr, 11, 11, n, d_cands.ptr, n, d_starts.ptr, s)
            ctx.fit_device_start(d_frames.ptr, 11, 11, n, d_cands.ptr, d_starts.ptr, n, d_first.ptr, s)
            ctx.fit_device_netinit(net, d_frames.ptr, 11, 11, n, d_cands.ptr, n, d_net.ptr, d_work.ptr, 3.0,
                                   stream=s)
        plain, first, final, work = d_plain.numpy(), d_first.numpy(), d_net.numpy(), d_work.numpy()
    refit = work["width"] == 0.0
    assert np.all((work["width"] == 0.0) | (work["width"] == -1.0))
    # the net-init result is the net-start fit, refitted from the reference start where flagged
    keep = ~refit
    assert np.array_equal(final["params"][keep], first["params"][keep])
    t_plain = plain["trials"].mean()
    t_net = (first["trials"].sum() + final["trials"][refit].sum()) / n
    acc = plain["status"] == 0
    both = acc & (final["status"] == 0)
    dxy = np.abs(final["params"][both][:, 2:4] - plain["params"][both][:, 2:4]).max(axis=1)
    _say(f"trials per ROI: plain {t_plain:.2f} (accepted {plain['trials'][acc].mean():.2f}), net-init {t_net:.2f} "
         f"(accepted {first['trials'][acc].mean():.2f}), fallback {refit.mean():.4f}; accepted plain {acc.mean():.4f} "
         f"net-init {np.mean(final['status'] == 0):.4f}; xy within 1e-3 where both accept: {np.mean(dxy <= 1e-3):.5f}")
    assert first["trials"][acc].mean() < 0.8 * plain["trials"][acc].mean()
    assert t_net < 0.85 * t_plain
    assert np.sum(acc & (final["status"] != 0)) <= 0.01 * acc.sum()
    assert np.mean(dxy <= 1e-3) >= 0.98
