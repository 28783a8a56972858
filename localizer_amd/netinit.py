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


def best_threshold(intensity, counted, seconds):
    """The LZR_OPT_NET_MAX_INTENSITY value, among measured candidates, with the highest rate
    of counted localizations: `intensity` holds the thresholds tried, `counted[k]` and
    `seconds[k]` what threshold k gave.  Ties go to the lower threshold (fewer network
    starts).  Returns (threshold, index)."""
    rate = np.asarray(counted, dtype=np.float64) / np.asarray(seconds, dtype=np.float64)
    k = int(np.flatnonzero(rate == rate.max())[0])
    return float(intensity[k]), k


def calibrate_selection(net, workload, nrois=1 << 21, stream_id=None, fractions=None, reps=3, log=None):
    """Chooses which candidates the network starts (LZR_OPT_NET_MAX_INTENSITY): on `nrois`
    fresh ROIs of the workload, net-initialized fits with the network on the dimmest fraction
    f of the candidates (by start intensity; f = 0 is the plain fit, f = 1 every candidate)
    are timed, and the threshold with the most localizations per second that agree with the
    plain fit (accepted by both, x and y within 1e-3 px) is returned with the table.  The
    network pays where the reference start leaves a long fit, i.e. on dim candidates; on bright
    ones it costs more than the few trials it saves."""
    stream_id = workload.extra.get("calibrate_stream", 45) if stream_id is None else stream_id
    fractions = (0.0, 0.05, 0.1, 0.15, 0.2, 0.3, 0.5, 1.0) if fractions is None else fractions
    rois, _, cands = simulate(workload, nrois, stream_id)
    n = len(cands)
    d_frames, d_cands = device.upload(rois), device.upload(cands)
    d_plain, d_out = device.empty(abi.FIT_DTYPE, n), device.empty(abi.FIT_DTYPE, n)
    d_work = device.empty(abi.START_DTYPE, n)
    s = device.stream_handle()
    w = 2 * abi.NET_HALF_WIDTH + 1
    order = np.sort(cands["intensity"])
    table = []
    with native.Context(abi.MODEL_SYMMETRIC, workload.initial_psf_width, 1.0, abi.NET_HALF_WIDTH) as ctx:
        ctx.fit_device(d_frames.ptr, w, w, n, d_cands.ptr, n, d_plain.ptr, s)
        device.synchronize()
        plain = d_plain.numpy()
        acc = plain["status"] == abi.ACCEPTED
        for f in fractions:
            if f <= 0.0:
                thr = -math.inf
            elif f >= 1.0:
                thr = math.inf
            else:
                thr = float(order[min(n - 1, int(f * n))])
            ctx.set_option(abi.OPT_NET_MAX_INTENSITY, thr)

            def call():
                if f <= 0.0:
                    ctx.fit_device(d_frames.ptr, w, w, n, d_cands.ptr, n, d_out.ptr, s)
                else:
                    ctx.fit_device_netinit(net, d_frames.ptr, w, w, n, d_cands.ptr, n, d_out.ptr, d_work.ptr, 3.0,
                                           stream=s)
            call()
            device.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                call()
            device.synchronize()
            sec = (time.perf_counter() - t0) / reps
            out = d_out.numpy()
            both = acc & (out["status"] == abi.ACCEPTED)
            dxy = np.abs(out["params"][:, 2:4] - plain["params"][:, 2:4]).max(axis=1)
            counted = int(np.sum(both & (dxy <= 1e-3)))
            table.append(dict(fraction=float(f), threshold=thr, seconds=sec, counted=counted))
            if log:
                log(f"net selection: dimmest {f:.0%} (intensity < {thr:.4g}): {sec * 1e3:.2f} ms, "
                    f"{counted} counted, {counted / sec:.4g} loc/s")
    thr, k = best_threshold([t["threshold"] for t in table], [t["counted"] for t in table],
                            [t["seconds"] for t in table])
    return dict(threshold=thr, fraction=table[k]["fraction"], rois=n, table=table)
