"""Training and use of the learned LM initializer (lzr_net_* of include/localizer_hip.h).

The network is trained on the GPU by the library's own forward / backward kernels, on
ROIs simulated by the fit's forward model (localizer_amd.synth) and labelled by the
plain LM fit (centroid start) of the same ROIs on the GPU.
"""
import math
import time
from dataclasses import replace

import numpy as np

from . import abi, device, native, synth


def simulate(workload, nrois, stream_id):
    """Independent 11 x 11 ROIs of the workload from Philox stream `stream_id`."""
    w = replace(workload, config_id=int(stream_id))
    return synth.generate_rois(w, int(nrois), abi.NET_HALF_WIDTH, float(workload.extra.get("max_offset", 1.0)))


def label(ctx, d_frames, nframes, d_cands, ncands):
    """Plain LM fits of ROI stacks (nx = ny = 11) on the GPU."""
    d_fits = device.empty(abi.FIT_DTYPE, ncands)
    n = 2 * abi.NET_HALF_WIDTH + 1
    ctx.fit_device(d_frames.ptr, n, n, nframes, d_cands.ptr, ncands, d_fits.ptr, device.stream_handle())
    return d_fits


def train(net, workload, nrois, epochs=1.0, batch=2048, lr=2e-3, lr_final=5e-5, stream_id=None, log=None):
    """Trains `net` on `nrois` fresh simulated ROIs; returns a dict of training statistics."""
    stream_id = workload.extra.get("train_stream", 4) if stream_id is None else stream_id
    t0 = time.time()
    rois, _, cands = simulate(workload, nrois, stream_id)
    t_sim = time.time() - t0
    n = 2 * abi.NET_HALF_WIDTH + 1
    d_frames = device.upload(rois)
    d_cands = device.upload(cands)
    with native.Context(abi.MODEL_SYMMETRIC, workload.initial_psf_width,
                        default_half_width=abi.NET_HALF_WIDTH) as ctx:
        d_fits = label(ctx, d_frames, nrois, d_cands, nrois)
    device.synchronize()
    fits = d_fits.numpy()
    used = int(np.sum((fits["status"] == abi.ACCEPTED)))
    steps = max(1, int(math.ceil(epochs * nrois / batch)))
    rng = np.random.default_rng(int(stream_id))
    losses = []
    batch = min(int(batch), nrois)
    k = nrois  # first step shuffles
    t1 = time.time()
    for s in range(steps):
        if k + batch > nrois:
            perm = rng.permutation(nrois)
            pc = d_cands.rows(perm)
            pf = d_fits.rows(perm)
            k = 0
        b = batch
        frac = s / max(1, steps - 1)
        cur = lr_final + 0.5 * (lr - lr_final) * (1.0 + math.cos(math.pi * frac))
        want = (s % 50 == 0) or s == steps - 1
        loss, _ = net.train_step_device(d_frames.ptr, n, n, nrois, pc.ptr + k * abi.CANDIDATE_DTYPE.itemsize,
                                        pf.ptr + k * abi.FIT_DTYPE.itemsize, b, cur, apply=True, want_loss=want,
                                        stream=device.stream_handle())
        k += b
        if want:
            losses.append((s, loss))
            if log:
                log(f"step {s}/{steps} lr {cur:.2e} loss {loss:.5f}")
    device.synchronize()
    return dict(steps=steps, batch=batch, rois=nrois, labelled_accepted=used, losses=losses,
                simulate_s=t_sim, train_s=time.time() - t1)
