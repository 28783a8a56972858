"""Benchmark: localizations/sec of the plain LM fit on one MI355X.

Workload (default): 10,000 frames of 256x256 uint16, 50 emitters/frame (500k ROIs),
sigma uniform 1.1-1.6 px, photons log-uniform 200-5000, background uniform 50-150 ADU,
Poisson noise, 9x9 windows, symmetric 5-parameter LM (initialPSFWidth 1.35).
One step = one fit of all 500k candidates with frames and candidates resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line.  The CPU baseline is the FP64 oracle (frame-parallel thread pool,
one frame per task, as the reference's PALMAnalysisController) on a bounded sample of
the same workload.

The line's "net_init" object is the second half of the metric, on the learned-initializer
workload: a CNN trained on the GPU (4M simulated 11x11 ROIs labelled by the plain LM)
predicts LM start points for 4M fresh ROIs; plain LM and net-initialized LM (with the
residual fallback) are timed on the same ROIs, and net-initialized localizations are
counted only where the plain LM also accepts and x, y agree within 1e-3 px.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from localizer_amd import abi, device, native, netinit, synth  # noqa: E402

METRIC = "localizations/sec on 1 MI355X: plain LM vs net-initialized LM at 1e-3 px"
PEAK_FP64_TFLOPS = 78.6  # the fit's passes and LM logic run in FP64 (vector FP64 peak)
PEAK_HBM_GBS = 8000.0
# HBM bytes per fitting call on the default workload (lm_persistent<5> two-wave one-round build,
# lm_straggler<5> and the epilogue), from rocprofv3 FETCH_SIZE / WRITE_SIZE, each scaled by
# the reading of a plain 512 MiB copy kernel (fetch x2.0, write x1.0).
MEASURED_TRAFFIC_BYTES = {"plain-lm-256px-10000frames-9x9": 1206.4e6}
# Algorithmic work per ROI (symmetric model): W = N[(1+T) F_f + (1+A)(F_J + F_NE)] + T F_solve
F_F, F_J, F_NE, F_SOLVE = 12.0, 6.0, 40.0, 300.0
F_F7, F_J7, F_NE7, F_SOLVE7 = 16.0, 30.0, 70.0, 700.0


def algorithmic_flops(fits, cands, default_h):
    h = np.where(cands["half_width"] > 0, cands["half_width"], default_h).astype(np.float64)
    N = (2 * h + 1) ** 2
    T = fits["trials"].astype(np.float64)
    A = fits["accepted_steps"].astype(np.float64)
    started = fits["iterations"] > 0
    sym = fits["model"] == abi.MODEL_SYMMETRIC
    w = np.where(sym, N * ((1 + T) * F_F + (1 + A) * (F_J + F_NE)) + T * F_SOLVE,
                 N * ((1 + T) * F_F7 + (1 + A) * (F_J7 + F_NE7)) + T * F_SOLVE7)
    return float(np.sum(w[started]))


def algorithmic_bytes(fits, cands, default_h):
    """The frozen per-ROI byte count: uint16 window pixels 2N + a 16-byte candidate (x, y, A0,
    b0 in FP32) + a 4(2p + 3)-byte result (p parameters, p deviations, 3 more fields in FP32):
    230 B for a 9x9 symmetric ROI."""
    h = np.where(cands["half_width"] > 0, cands["half_width"], default_h).astype(np.float64)
    N = (2 * h + 1) ** 2
    p = np.where(fits["model"] == abi.MODEL_SYMMETRIC, 5, np.where(fits["model"] == abi.MODEL_FIXED_WIDTH, 4, 7))
    return float(np.sum(2 * N + 16 + 4 * (2 * p + 3)))


def record_bytes(fits, cands, default_h):
    """What this engine's records move per ROI: window pixels + its 40-byte candidate + its
    192-byte result record."""
    h = np.where(cands["half_width"] > 0, cands["half_width"], default_h).astype(np.float64)
    N = (2 * h + 1) ** 2
    return float(np.sum(2 * N + abi.CANDIDATE_DTYPE.itemsize + abi.FIT_DTYPE.itemsize))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, frames, cands, target_seconds, nthreads):
    """Oracle throughput on a bounded prefix of the workload (whole frames), best of 3."""
    from oracle import pyoracle as po

    def run(nf):
        sel = cands["frame"] < nf
        t0 = time.perf_counter()
        out = po.fit_batch(frames[:nf], cands[sel], abi.MODEL_SYMMETRIC, w.initial_psf_width,
                           default_half_width=w.half_width[0], nthreads=nthreads)
        dt = time.perf_counter() - t0
        return int(np.sum(out["status"] == 0)), int(sel.sum()), dt

    nf = min(50, w.nframes)
    acc, n, dt = run(nf)
    if dt < target_seconds / 3 and nf < w.nframes:
        nf = int(min(w.nframes, max(nf, nf * target_seconds / 3 / max(dt, 1e-3))))
    runs = [run(nf) for _ in range(3)]
    acc, n, dt = min(runs, key=lambda r: r[2])
    return {
        "value": acc / dt,
        "unit": "localizations/s",
        "cores": nthreads,
        "kind": "port",
        "sample": f"{nf} of {w.nframes} frames ({n} ROIs, {acc} accepted), best of 3 runs "
        f"({', '.join(f'{r[2]:.2f}' for r in runs)} s): the FP64 oracle (oracle/lzr_oracle.c, gcc -O2, "
        f"no fast-math; the reference's own fitter needs libgsl, absent) on {nthreads} threads of a "
        f"{cpu_model()}, one frame per task as the reference's PALMAnalysisController",
    }


def parity_vs_oracle(w, frames, cands, fits, nthreads):
    """Per-parameter agreement of the timed GPU fits with the FP64 oracle over the whole
    workload (tests/parity.py metrics; the bar: x, y within 1e-3 px, the rest within 1e-4)."""
    from oracle import pyoracle as po
    from tests.parity import compare

    t0 = time.perf_counter()
    ref = po.fit_batch(frames, cands, abi.MODEL_SYMMETRIC, w.initial_psf_width, default_half_width=w.half_width[0],
                       nthreads=nthreads)
    st = compare(ref, fits, abi.MODEL_SYMMETRIC)
    keep = ("status_agreement", "both", "xy_max_px", "xy_frac_within_1e-3", "conv_n", "conv_xy_max_px",
            "conv_xy_frac_within_1e-3")
    out = {k: st[k] for k in keep if k in st}
    for nm in ("A", "r", "b"):
        for pre in ("", "conv_"):
            for suf in ("frac_within_1e-4", "max_rel", "max_abs"):
                k = f"{pre}{nm}_{suf}"
                if k in st:
                    out[k] = st[k]
    for nm in ("A", "r", "x0", "y0", "b"):
        out[f"dev_{nm}_frac_within_1e-4"] = st[f"dev_{nm}_frac_within_1e-4"]
    for f in ("chi", "integral", "integral_deviation"):
        out[f"{f}_frac_within_1e-4"] = st[f"{f}_frac_within_1e-4"]
    out["oracle_seconds"] = time.perf_counter() - t0
    return out


def pcie_inclusive(ctx, frames, cands, reps=3):
    """lzr_fit_batch on host buffers: frames and candidates copied in, results copied out
    (pageable memory), per call.  Not the headline value (inputs there are resident in HBM)."""
    ctx.fit(frames[:1], cands[cands["frame"] < 1])  # staging allocated
    best = 1e30
    for _ in range(reps):
        t0 = time.perf_counter()
        res = ctx.fit(frames, cands)
        best = min(best, time.perf_counter() - t0)
    acc = int(np.sum(res["status"] == 0))
    return {"value": acc / best, "unit": "localizations/s", "ms_per_call": best * 1e3,
            "bytes_in": int(frames.nbytes + cands.nbytes), "bytes_out": int(res.nbytes),
            "entry": "lzr_fit_batch (host buffers, pageable)", "best_of": reps}


NET_FLOPS_PER_ROI = 2.0 * (81 * 32 * 9 + 49 * 32 * 288 + 1568 * 64 + 64 * 5)
PEAK_BF16_TFLOPS = 2500.0


def controller_leg(ctx, w, frames, cands, fits, log, thread_counts=(16, 64, 256), inflight=4, direct_frames=400):
    """The reference controller's call pattern: nthreads workers sharing one fitter, one call
    per frame (source/PALMAnalysis.cpp:204-252, call at :234), host buffers in and out,
    through the host batcher (lzr_batcher_fit_frame); beside it the same pattern with one
    lzr_fit_batch call per frame."""
    import threading

    order = np.argsort(cands["frame"], kind="stable")
    sc = cands[order]
    bounds = np.searchsorted(sc["frame"], np.arange(w.nframes + 1))
    lists = []
    for f in range(w.nframes):
        c = sc[bounds[f]:bounds[f + 1]].copy()
        c["frame"] = 0
        lists.append(c)

    def run(fn, nf):
        out, errs = [None] * nf, []

        def work(t):
            try:
                for f in range(t, nf, nthreads):
                    out[f] = fn(f)
            except Exception as e:  # noqa: BLE001 -- reported
                errs.append(repr(e))

        th = [threading.Thread(target=work, args=(t,)) for t in range(nthreads)]
        t0 = time.perf_counter()
        for t in th:
            t.start()
        for t in th:
            t.join()
        if errs:
            raise RuntimeError(errs[0])
        return time.perf_counter() - t0, out

    want = fits[order].tobytes()
    sweep = []
    for nt in thread_counts:
        nthreads = nt
        with native.Batcher(ctx, w.nx, w.ny, max_frames=256, max_rois=1 << 15, max_wait_us=500,
                            inflight=inflight) as b:
            run(lambda f: b.fit_frame(frames[f], lists[f]), min(256, w.nframes))  # warm-up
            st0 = b.stats()
            sec, out = run(lambda f: b.fit_frame(frames[f], lists[f]), w.nframes)
            st1 = b.stats()
        got = np.concatenate(out)
        acc = int(np.sum(got["status"] == 0))
        nb = st1["batches"] - st0["batches"]
        sweep.append({"threads": nt, "seconds": sec, "frames_per_s": w.nframes / sec, "localizations_per_s": acc / sec,
                      "batches": nb, "mean_frames_per_batch": w.nframes / max(nb, 1),
                      "records_identical_to_device_call": bool(got.tobytes() == want)})
        log(f"controller pattern ({nt} threads, one call per frame, {inflight} batches in flight): "
            f"{sweep[-1]['frames_per_s']:.0f} frames/s ({sweep[-1]['localizations_per_s']:.4g} loc/s, "
            f"{sweep[-1]['mean_frames_per_batch']:.1f} frames per batch, identical "
            f"{sweep[-1]['records_identical_to_device_call']})")
    nthreads = thread_counts[0]
    nd = min(direct_frames, w.nframes)
    sec_d, _ = run(lambda f: ctx.fit(frames[f], lists[f]), nd)
    log(f"one lzr_fit_batch call per frame ({nthreads} threads): {nd / sec_d:.0f} frames/s")
    return {"frames": w.nframes, "inflight": inflight, "batcher": sweep,
            "one_call_per_frame": {"threads": nthreads, "frames": nd, "frames_per_s": nd / sec_d,
                                   "entry": "lzr_fit_batch per frame (nframes 1), calls serialised on the context"},
            "entry": "lzr_batcher_fit_frame from Python threads (host uint16 frame in, records out)"}


def netinit_vs_oracle(W, frames, cands, plain, net, m):
    """Plain and net-init records against the FP64 oracle (reference start) on the first m
    ROIs: accepted/rejected agreement, x/y within 1e-3 px and A, r, b within 1e-4 relative
    where both accept."""
    from oracle import pyoracle as po

    nthreads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    t0 = time.perf_counter()
    ref = po.fit_batch(frames[:m], cands[:m], abi.MODEL_SYMMETRIC, W.initial_psf_width,
                       default_half_width=abi.NET_HALF_WIDTH, nthreads=nthreads)
    out = {"rois": int(m), "oracle_seconds": time.perf_counter() - t0}
    racc = ref["status"] == 0
    for name, fits in (("plain", plain[:m]), ("netinit", net[:m])):
        acc = fits["status"] == 0
        both = acc & racc
        d = np.abs(fits["params"][both] - ref["params"][both])
        rel = d / np.maximum(np.abs(ref["params"][both]), 1e-300)
        out[name] = {"accepted_agreement": float(np.mean(acc == racc)),
                     "status_agreement": float(np.mean(fits["status"] == ref["status"])),
                     "both_accepted": int(both.sum()),
                     "xy_within_1e-3": float(np.mean(d[:, 2:4].max(1) <= 1e-3)),
                     "xy_max_px": float(d[:, 2:4].max()) if both.any() else 0.0,
                     "A_r_b_within_1e-4": [float(np.mean(rel[:, k] <= 1e-4)) for k in (0, 1, 4)]}
    return out


def net_init_leg(train_rois, eval_rois, epochs, steps, warmup, log, oracle_rois=100000):
    """Plain vs net-initialized LM on the learned-initializer workload (one GPU)."""
    import torch

    W = synth.LEARNED_INIT_11
    net = native.Net(W.initial_psf_width, seed=3)
    tr = netinit.train(net, W, train_rois, epochs=epochs, batch=2048, lr=2e-3)
    log(f"net: trained on {train_rois} ROIs, {tr['steps']} steps in {tr['train_s']:.2f} s "
        f"(simulation {tr['simulate_s']:.1f} s), last loss {tr['losses'][-1][1]:.4f}")
    # which candidates the network starts: chosen on an independent calibration set
    sel = netinit.calibrate_selection(net, W, nrois=min(1 << 21, eval_rois), log=log)
    log(f"net selection: network on candidates with intensity < {sel['threshold']:.4g} "
        f"(the dimmest {sel['fraction']:.0%} of the calibration set)")
    frames, _, cands = netinit.simulate(W, eval_rois, W.extra["eval_stream"])
    n = len(cands)
    d_frames, d_cands = device.upload(frames), device.upload(cands)
    d_plain = device.empty(abi.FIT_DTYPE, n)
    d_net = device.empty(abi.FIT_DTYPE, n)
    d_work = device.empty(abi.START_DTYPE, n)
    d_starts = device.empty(abi.START_DTYPE, n)
    d_nrefit = device.empty(np.uint32, 1)
    ctx = native.Context(abi.MODEL_SYMMETRIC, W.initial_psf_width, 1.0, abi.NET_HALF_WIDTH)
    ctx.set_option(abi.OPT_NET_MAX_INTENSITY, sel["threshold"])
    ctx_all = native.Context(abi.MODEL_SYMMETRIC, W.initial_psf_width, 1.0, abi.NET_HALF_WIDTH)
    stream = torch.cuda.current_stream()
    s = stream.cuda_stream

    def plain():
        ctx.fit_device(d_frames.ptr, 11, 11, n, d_cands.ptr, n, d_plain.ptr, s)

    def netinit_fit():
        ctx.fit_device_netinit(net, d_frames.ptr, 11, 11, n, d_cands.ptr, n, d_net.ptr, d_work.ptr, 3.0,
                               d_nrefit=d_nrefit.ptr, stream=s)

    d_all = device.empty(abi.FIT_DTYPE, n)
    d_work_all = device.empty(abi.START_DTYPE, n)

    def netinit_all():
        ctx_all.fit_device_netinit(net, d_frames.ptr, 11, 11, n, d_cands.ptr, n, d_all.ptr, d_work_all.ptr, 3.0,
                                   stream=s)

    d_strict = device.empty(abi.FIT_DTYPE, n)
    d_work2 = device.empty(abi.START_DTYPE, n)

    def netinit_strict():
        ctx.fit_device_netinit(net, d_frames.ptr, 11, 11, n, d_cands.ptr, n, d_strict.ptr, d_work2.ptr, 3.0,
                               refit_rejected=True, stream=s)

    def predict():
        net.predict_device(d_frames.ptr, 11, 11, n, d_cands.ptr, n, d_starts.ptr, s)

    def timed(fn):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        for a, b in ev:
            a.record(stream)
            fn()
            b.record(stream)
        torch.cuda.synchronize()
        return float(np.mean([a.elapsed_time(b) for a, b in ev]))

    ms_plain = timed(plain)
    ms_net = timed(netinit_fit)
    ms_all = timed(netinit_all)
    ms_pred = timed(predict)
    ms_strict = timed(netinit_strict)
    P, F, wk = d_plain.numpy(), d_net.numpy(), d_work.numpy()
    S = d_strict.numpy()
    acc = P["status"] == 0

    def counted(R):
        dxy = np.abs(R["params"][:, 2:4] - P["params"][:, 2:4]).max(axis=1)
        return acc & (R["status"] == 0) & (dxy <= 1e-3), dxy

    good, dxy = counted(F)
    both = acc & (F["status"] == 0)
    good_all, _ = counted(d_all.numpy())
    refit = wk["width"] == 0.0
    out = {
        "workload": W.name,
        "rois": n,
        "window": "11x11",
        "initial_psf_width": W.initial_psf_width,
        "plain_value": float(acc.sum() / (ms_plain * 1e-3)),
        "value": float(good.sum() / (ms_net * 1e-3)),
        "speedup": float(good.sum() / (ms_net * 1e-3)) / float(acc.sum() / (ms_plain * 1e-3)),
        "ms_plain": ms_plain,
        "ms_netinit": ms_net,
        "ms_predict": ms_pred,
        "accepted_plain": int(acc.sum()),
        "accepted_netinit": int(np.sum(F["status"] == 0)),
        "counted_netinit": int(good.sum()),
        "xy_within_1e-3": float(np.mean(dxy[both] <= 1e-3)),
        "status_agreement": float(np.mean(F["status"] == P["status"])),
        "selection": {"rule": "the network starts candidates with intensity below the threshold, the others "
                              "start from the reference start; the threshold maximizes counted localizations/s "
                              "on an independent calibration set",
                      "threshold": sel["threshold"], "calibration_fraction": sel["fraction"],
                      "fraction": float(np.mean(cands["intensity"] < sel["threshold"])),
                      "calibration": sel["table"]},
        "all_rois": {"ms": ms_all, "value": float(good_all.sum() / (ms_all * 1e-3)),
                     "speedup": float(good_all.sum() / (ms_all * 1e-3)) / float(acc.sum() / (ms_plain * 1e-3))},
        "fallback_fraction": float(refit.mean()),
        "fallback_counted_on_device": int(d_nrefit.numpy()[0]),
        "fallback_rule": "refit from the reference start when the net-start fit ends with |f| above 3x |f| at "
                         "the reference start",
        "accepted_agreement_with_plain": float(np.mean((F["status"] == 0) == acc)),
        "trials_plain": float(P["trials"].mean()),
        "trials_netinit_final": float(F["trials"].mean()),
        "refit_rejected_policy": {
            "ms": ms_strict, "fallback_fraction": float(np.mean(d_work2.numpy()["width"] == 0.0)),
            "accepted_agreement_with_plain": float(np.mean((S["status"] == 0) == acc)),
            "status_agreement": float(np.mean(S["status"] == P["status"])),
            "rule": "also refit net-start fits that were rejected or did not converge"},
        "vs_fp64_oracle": netinit_vs_oracle(W, frames, cands, P, F, oracle_rois),
        "train": {"rois": train_rois, "steps": tr["steps"], "batch": tr["batch"], "seconds": tr["train_s"],
                  "last_loss": tr["losses"][-1][1]},
        "network": {"bound": "mfma", "achieved": NET_FLOPS_PER_ROI * n / (ms_pred * 1e-3) / 1e12,
                    "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                    "frac": NET_FLOPS_PER_ROI * n / (ms_pred * 1e-3) / 1e12 / PEAK_BF16_TFLOPS,
                    "note": "algorithmic FLOPs of the CNN forward / duration of lzr_net_predict_device "
                            "(patch prep + forward + decode)"},
    }
    log(f"net-init: plain {ms_plain:.2f} ms ({out['plain_value']:.4g} loc/s), net-init {ms_net:.2f} ms "
        f"({out['value']:.4g} loc/s counted), predict {ms_pred:.2f} ms, speedup {out['speedup']:.3f}; "
        f"network on every candidate {ms_all:.2f} ms, speedup {out['all_rois']['speedup']:.3f}")
    ctx.close()
    ctx_all.close()
    net.close()
    return out


def mixed_leg(steps, warmup, log, frames_per_block=1000, parity=True):
    """Batched uneven workload: 2M ROIs, windows 5x5..15x15, 60 % symmetric / 40 %
    ellipsoidal; one call over all ROIs (default and autotuned launch configuration) and
    frame blocks replayed through one captured HIP graph."""
    import torch

    w = synth.MIXED_2M
    nthreads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    t0 = time.perf_counter()
    frames, _, cands = synth.generate(w, nthreads=nthreads)
    log(f"mixed: generated {w.nframes} frames, {cands.size} ROIs in {time.perf_counter() - t0:.1f} s")
    n = cands.size
    d_frames, d_cands = device.upload(frames), device.upload(cands)
    if not parity:
        del frames
    d_out = device.empty(abi.FIT_DTYPE, n)
    stream = torch.cuda.current_stream()
    s = stream.cuda_stream
    ctx = native.Context(abi.MODEL_SYMMETRIC, w.initial_psf_width)
    ctx.set_option(abi.OPT_MAX_HALF_WIDTH, w.half_width[1])

    def timed(fn):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        for a, b in ev:
            a.record(stream)
            fn()
            b.record(stream)
        torch.cuda.synchronize()
        return float(np.mean([a.elapsed_time(b) for a, b in ev]))

    def one_call():
        ctx.fit_device(d_frames.ptr, w.nx, w.ny, w.nframes, d_cands.ptr, n, d_out.ptr, s)

    ms_default = timed(one_call)
    fits = d_out.numpy()
    accepted = int(np.sum(fits["status"] == 0))
    # autotune on the batch itself (one launch per configuration; the hand-off count, and so the
    # straggler group size, is a property of the batch), then time it again
    n_tune = n
    best = ctx.autotune_device(d_frames.ptr, w.nx, w.ny, w.nframes, d_cands.ptr, n_tune, d_out.ptr, s)
    chosen = {"occupancy": ctx.get_option(abi.OPT_OCCUPANCY), "waves_per_cu": ctx.get_option(abi.OPT_WAVES_PER_CU),
              "rounds": ctx.get_option(abi.OPT_ROUNDS), "straggler_trials": ctx.get_option(abi.OPT_STRAGGLER_TRIALS),
              "straggler_group": ctx.get_option(abi.OPT_STRAGGLER_GROUP)}
    ms_tuned = timed(one_call)
    # frame blocks through captured graphs: candidates of block b are staged (frames stay
    # resident), a graph replays, results are copied out.  Blocks rotate over `pipes`
    # pipelines (context + graph + stream), so one block's slowest fits overlap the next
    # blocks' work.
    per_block = w.emitters_per_frame * frames_per_block
    nblocks = (n + per_block - 1) // per_block
    counts = torch.tensor([min(per_block, n - b * per_block) for b in range(nblocks)], dtype=torch.int32,
                          device="cuda")
    cb, fb = abi.CANDIDATE_DTYPE.itemsize, abi.FIT_DTYPE.itemsize
    cview, oview = d_cands.t, d_out.t
    pipes = []
    for _ in range(4):
        pc = native.Context(abi.MODEL_SYMMETRIC, w.initial_psf_width)
        pc.set_option(abi.OPT_MAX_HALF_WIDTH, w.half_width[1])
        st_c = device.empty(abi.CANDIDATE_DTYPE, per_block)
        st_o = device.empty(abi.FIT_DTYPE, per_block)
        st_n = torch.zeros(1, dtype=torch.int32, device="cuda")
        g = pc.graph(d_frames.ptr, w.nx, w.ny, w.nframes, st_c.ptr, st_n.data_ptr(), per_block, st_o.ptr)
        pipes.append((pc, st_c, st_o, st_n, g, torch.cuda.Stream()))

    def blocks():
        start_ev = torch.cuda.Event()
        start_ev.record(stream)
        for _, _, _, _, _, ps in pipes:
            ps.wait_event(start_ev)
        for b in range(nblocks):
            pc, st_c, st_o, st_n, g, ps = pipes[b % len(pipes)]
            k = int(min(per_block, n - b * per_block))
            with torch.cuda.stream(ps):
                st_c.t[: k * cb].copy_(cview[b * per_block * cb:(b * per_block + k) * cb])
                st_n.copy_(counts[b:b + 1])
                g.launch(ps.cuda_stream)
                oview[b * per_block * fb:(b * per_block + k) * fb].copy_(st_o.t[: k * fb])
        for *_, ps in pipes:
            stream.wait_stream(ps)

    ms_graph = timed(blocks)
    gfits = d_out.numpy()
    # a fit's result depends on its spot only (straggler groups compute the one-lane pass bit
    # for bit): the blocks must reproduce the one call exactly
    status_diff = int(np.sum(gfits["status"] != fits["status"]))
    both = np.flatnonzero((gfits["status"] == 0) & (fits["status"] == 0))
    xi = np.where(cands["model"][both] == abi.MODEL_ELLIPSOIDAL, 3, 2)
    dxy = np.zeros(both.size)
    for off in (0, 1):
        dxy = np.maximum(dxy, np.abs(gfits["params"][both, xi + off] - fits["params"][both, xi + off]))
    dpos = float(np.max(dxy, initial=0.0))
    far = int(np.sum(dxy > 1e-3))
    for pc, *_rest, g, _ps in pipes:
        g.close()
        pc.close()
    ctx.close()
    out = {
        "workload": w.name,
        "rois": int(n),
        "windows": "5x5..15x15",
        "ellipsoidal_fraction": float(np.mean(cands["model"] == abi.MODEL_ELLIPSOIDAL)),
        "accepted": accepted,
        "value": accepted / (ms_tuned * 1e-3),
        "ms_one_call_default": ms_default,
        "ms_one_call_autotuned": ms_tuned,
        "autotune": {"sample_rois": int(n_tune), "best_sample_ms": best, "chosen": chosen},
        "graph_blocks": {"frames_per_block": frames_per_block, "blocks": int(nblocks), "pipelines": len(pipes),
                         "ms": ms_graph,
                         "value": accepted / (ms_graph * 1e-3), "status_differs": status_diff, "max_position_difference_px": dpos,
                         "positions_differing_over_1e-3_px": far},
        "trials_mean": float(fits["trials"].mean()),
    }
    if parity:
        # every one of the 2M one-call fits against the FP64 oracle, per model class
        from oracle import pyoracle as po
        from tests.parity import compare

        t0 = time.perf_counter()
        ref = po.fit_batch(frames, cands, abi.MODEL_SYMMETRIC, w.initial_psf_width, nthreads=nthreads)
        out["parity_vs_fp64_oracle"] = {"oracle_seconds": time.perf_counter() - t0}
        for m, name in ((abi.MODEL_SYMMETRIC, "symmetric"), (abi.MODEL_ELLIPSOIDAL, "ellipsoidal")):
            sel = fits["model"] == m
            st = compare(ref[sel], fits[sel], m)
            keep = {k: st[k] for k in ("n", "status_agreement", "both", "xy_frac_within_1e-3", "xy_max_px",
                                       "conv_n", "conv_xy_frac_within_1e-3") if k in st}
            for nm in ("A", "r", "sx", "sy", "rho", "b"):
                for pre in ("", "conv_"):
                    k = f"{pre}{nm}_frac_within_1e-4"
                    if k in st:
                        keep[k] = st[k]
            out["parity_vs_fp64_oracle"][name] = keep
        del frames
        log(f"mixed parity vs FP64 oracle: " + ", ".join(
            f"{k} status {v['status_agreement']:.6f} x/y within 1e-3 {v['xy_frac_within_1e-3']:.6f}"
            for k, v in out["parity_vs_fp64_oracle"].items() if isinstance(v, dict)))
    identical = bool(np.array_equal(gfits.view(np.uint8), fits.view(np.uint8)))
    out["graph_blocks"]["records_identical_to_one_call"] = identical
    log(f"mixed: one call {ms_default:.2f} ms, autotuned {ms_tuned:.2f} ms ({chosen}), graph blocks {ms_graph:.2f} ms "
        f"(records identical to the one call: {identical}; status differs for {status_diff} fits, positions by "
        f"<= {dpos:.1e} px)")
    return out


DENSE_PFA = 20.0


DENSE_PIPELINES = 3  # frame blocks in flight (one stream each)
RENDER_SCALE = 10.0
RENDER_DEVIATION = 0.15


def render_leg(d_fits, nfits, w, log, reps=5):
    """Localization bitmap of the plain leg's accepted fits: positions_from_fits on the
    device, then the bitmap at 10x (constant deviation 0.15 px, integral weighing)."""
    import torch

    d_pos = device.empty(abi.POSITION_DTYPE, nfits)
    d_count = device.empty(np.int32, 1)
    width, height = native.render_image_size(w.nx, w.ny, RENDER_SCALE)
    d_img = device.empty(np.float64, width * height)
    s = torch.cuda.current_stream()
    with native.Renderer() as r:
        r.positions_from_fits_device(d_fits.data_ptr(), nfits, d_pos.ptr, d_count.ptr, s.cuda_stream)
        n = int(d_count.numpy()[0])

        def once():
            r.render_device(d_pos.ptr, n, w.nx, w.ny, RENDER_SCALE, abi.DEVIATION_CONSTANT, [RENDER_DEVIATION],
                            abi.WEIGHING_INTEGRAL, d_img.ptr, s.cuda_stream)

        once()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            once()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / reps
    box = 2 * int(np.ceil(4 * RENDER_DEVIATION * (width - 1) / (w.nx - 1))) + 1
    pairs = n * box * box
    log(f"render: {n} positions -> {width}x{height} image in {ms:.3f} ms ({n / ms * 1e3:.3e} positions/s, "
        f"~{pairs / ms * 1e3:.3e} pixel-position terms/s)")
    return {"positions": n, "image": f"{width}x{height}", "scale": RENDER_SCALE, "deviation": RENDER_DEVIATION,
            "ms": round(ms, 3), "positions_per_s": n / ms * 1e3, "pixel_terms_per_s_approx": pairs / ms * 1e3}


def dense_leg(nframes, block_frames, latency_frames, log, pipelines=DENSE_PIPELINES):
    """Dense end-to-end stream: frames -> GLRT segmentation -> 8-way particle finding ->
    11x11 patch gather -> CNN start -> LM with residual fallback, on the device; frames
    resident in HBM, one host read of the candidate count per block."""
    import torch

    w = synth.DENSE_2048.scaled(nframes)
    nthreads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    t0 = time.perf_counter()
    frames, _, _ = synth.generate(w, nthreads=nthreads)
    log(f"dense: generated {w.nframes} frames of {w.nx}x{w.ny} in {time.perf_counter() - t0:.1f} s")
    d_frames = device.upload(frames)
    del frames
    npx = w.nx * w.ny
    net = native.Net(w.initial_psf_width, seed=5)
    tr = netinit.train(net, synth.DENSE_ROIS_11, 1 << 20, epochs=1.0, batch=2048, lr=2e-3)
    log(f"dense: network trained on {tr['rois']} ROIs in {tr['train_s']:.2f} s")
    sel = netinit.calibrate_selection(net, synth.DENSE_ROIS_11, nrois=1 << 20, log=log)
    log(f"dense: network on candidates with intensity < {sel['threshold']:.4g}")
    cap = block_frames * 2048
    status_off = abi.FIT_DTYPE.fields["status"][1]

    class Pipe:
        """One block in flight: its own finder, context, buffers and stream (the network
        is shared; its calls on different streams are ordered by the library)."""

        def __init__(self):
            self.finder = native.Finder(w.initial_psf_width, DENSE_PFA)
            self.ctx = native.Context(abi.MODEL_SYMMETRIC, w.initial_psf_width, 1.0, w.half_width[0])
            self.ctx.set_option(abi.OPT_NET_MAX_INTENSITY, sel["threshold"])
            self.d_cands = device.empty(abi.CANDIDATE_DTYPE, cap)
            self.d_fits = device.empty(abi.FIT_DTYPE, cap)
            self.d_work = device.empty(abi.START_DTYPE, cap)
            self.d_count = torch.zeros(1, dtype=torch.int32, device="cuda")
            self.stream = torch.cuda.Stream()
            self.acc = torch.zeros((), dtype=torch.int64, device="cuda")

        def run(self, f0, nb, netinit_fit):
            """Segmentation, then (after a read of this block's candidate count, which waits
            for this pipeline's stream only) the fit; accepted fits counted on the device."""
            s = self.stream.cuda_stream
            with torch.cuda.stream(self.stream):
                ptr = d_frames.ptr + f0 * npx * 2
                self.finder.find_device(ptr, w.nx, w.ny, nb, self.d_cands.ptr, cap, self.d_count.data_ptr(), s)
                n = int(self.d_count.item())
                if n < 0:
                    raise RuntimeError("dense stream: candidate capacity exceeded")
                if n:
                    if netinit_fit:
                        self.ctx.fit_device_netinit(net, ptr, w.nx, w.ny, nb, self.d_cands.ptr, n, self.d_fits.ptr,
                                                    self.d_work.ptr, 3.0, stream=s)
                    else:
                        self.ctx.fit_device(ptr, w.nx, w.ny, nb, self.d_cands.ptr, n, self.d_fits.ptr, s)
                    st = self.d_fits.t[: n * abi.FIT_DTYPE.itemsize].view(n, abi.FIT_DTYPE.itemsize)
                    st = st[:, status_off:status_off + 4].contiguous().view(torch.int32)
                    self.acc += (st == 0).sum()
            return n

        def close(self):
            self.finder.close()
            self.ctx.close()

    pipes = [Pipe() for _ in range(pipelines)]

    def stream_all(netinit_fit):
        for p in pipes:  # warm-up: every pipeline's buffers and work memory
            p.run(0, min(block_frames, nframes), netinit_fit)
        torch.cuda.synchronize()
        for p in pipes:
            p.acc.zero_()
        torch.cuda.synchronize()
        cands = 0
        t = time.perf_counter()
        for b, f0 in enumerate(range(0, nframes, block_frames)):
            cands += pipes[b % len(pipes)].run(f0, min(block_frames, nframes - f0), netinit_fit)
        torch.cuda.synchronize()
        sec = time.perf_counter() - t
        return sec, cands, int(sum(int(p.acc.item()) for p in pipes))

    # each mode twice, alternating, best of two: the first pass over the stream also warms
    # clocks and caches for whichever mode runs second
    runs = {True: [], False: []}
    for netinit_fit in (True, False, True, False):
        runs[netinit_fit].append(stream_all(netinit_fit))
    sec_net, cands, acc_net = min(runs[True], key=lambda r: r[0])
    sec_plain, _, acc_plain = min(runs[False], key=lambda r: r[0])
    lat = []
    for f in range(min(latency_frames, nframes)):
        torch.cuda.synchronize()
        t = time.perf_counter()
        pipes[0].run(f, 1, True)
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - t) * 1e3)
    lat = np.array(lat)
    out = {
        "workload": w.name,
        "frames": nframes,
        "frame_px": f"{w.nx}x{w.ny}",
        "glrt_pfa": DENSE_PFA,
        "window": f"{2 * w.half_width[0] + 1}x{2 * w.half_width[0] + 1}",
        "block_frames": block_frames,
        "pipelines": pipelines,
        "candidates_per_frame": cands / nframes,
        "netinit": {"frames_per_s": nframes / sec_net, "localizations_per_s": acc_net / sec_net,
                    "accepted": acc_net, "seconds": sec_net,
                    "network_below_intensity": sel["threshold"], "calibration_fraction": sel["fraction"]},
        "plain": {"frames_per_s": nframes / sec_plain, "localizations_per_s": acc_plain / sec_plain,
                  "accepted": acc_plain, "seconds": sec_plain},
        "latency_ms_per_frame": {"frames": int(lat.size), "p50": float(np.percentile(lat, 50)),
                                 "p99": float(np.percentile(lat, 99)), "mode": "one frame per call, netinit"},
    }
    log(f"dense: {cands / nframes:.1f} candidates/frame; net-init {nframes / sec_net:.1f} frames/s "
        f"({acc_net / sec_net:.4g} loc/s), plain {nframes / sec_plain:.1f} frames/s; latency p50 "
        f"{out['latency_ms_per_frame']['p50']:.2f} ms p99 {out['latency_ms_per_frame']['p99']:.2f} ms")
    for p in pipes:
        p.close()
    net.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--frames", type=int, default=0, help="override the frame count (testing)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--occupancy", type=int, default=0, help="fitting-kernel register budget (1 or 2 waves/SIMD)")
    ap.add_argument("--waves-per-cu", type=int, default=0, help="persistent grid size per CU (0: resident max)")
    ap.add_argument("--rounds", type=int, default=0, help="fitting-kernel rounds (fits parked in between)")
    ap.add_argument("--park-threshold", type=int, default=-1, help="busy lanes at which a wave parks its fits")
    ap.add_argument("--straggler-trials", type=int, default=-1,
                    help="trials after which a fit moves to the 8-lane straggler kernel (0: off)")
    ap.add_argument("--straggler-group", type=int, default=-1, help="lanes per handed-off fit (0: by count)")
    ap.add_argument("--no-net", action="store_true", help="skip the net-initialized LM leg")
    ap.add_argument("--only-net", action="store_true", help="run only the net-initialized LM leg (profiling)")
    ap.add_argument("--no-mixed", action="store_true", help="skip the batched mixed-workload leg")
    ap.add_argument("--block-frames", type=int, default=10000, help="frames per graph-replayed block (mixed leg)")
    ap.add_argument("--no-dense", action="store_true", help="skip the dense end-to-end stream leg")
    ap.add_argument("--no-render", action="store_true", help="skip the localization-bitmap leg")
    ap.add_argument("--no-controller", action="store_true", help="skip the per-frame (controller pattern) leg")
    ap.add_argument("--dense-frames", type=int, default=1000)
    ap.add_argument("--dense-block", type=int, default=50, help="frames per block of the dense stream")
    ap.add_argument("--dense-pipelines", type=int, default=DENSE_PIPELINES, help="frame blocks in flight")
    ap.add_argument("--net-train-rois", type=int, default=4 << 20)
    ap.add_argument("--net-eval-rois", type=int, default=4000000)
    ap.add_argument("--net-epochs", type=float, default=8.0)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device once (scales rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch

        n = args.calibrate_copy << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": n}), flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import torch

    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.only_net:
        net = net_init_leg(args.net_train_rois, args.net_eval_rois, args.net_epochs, max(1, args.steps),
                           args.warmup, lambda m: print(m, file=sys.stderr, flush=True))
        print(json.dumps({"net_init": net}), flush=True)
        return
    w = synth.PLAIN_GPU_256
    if args.frames:
        w = w.scaled(args.frames)
    nthreads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    t0 = time.perf_counter()
    frames, truth, cands = synth.generate(w, nthreads=nthreads)
    print(f"generated {w.nframes} frames, {cands.size} candidates in {time.perf_counter() - t0:.1f} s",
          file=sys.stderr, flush=True)

    dev = torch.device("cuda:0")
    d_frames = torch.from_numpy(frames.view(np.int16)).to(dev)
    d_cands = torch.from_numpy(cands.view(np.uint8)).to(dev)
    d_out = torch.zeros(cands.size * abi.FIT_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()
    ctx = native.Context(abi.MODEL_SYMMETRIC, w.initial_psf_width, 1.0, w.half_width[0])
    if args.occupancy:
        ctx.set_option(abi.OPT_OCCUPANCY, args.occupancy)
    if args.waves_per_cu:
        ctx.set_option(abi.OPT_WAVES_PER_CU, args.waves_per_cu)
    if args.rounds:
        ctx.set_option(abi.OPT_ROUNDS, args.rounds)
    if args.park_threshold >= 0:
        ctx.set_option(abi.OPT_PARK_THRESHOLD, args.park_threshold)
    if args.straggler_trials >= 0:
        ctx.set_option(abi.OPT_STRAGGLER_TRIALS, args.straggler_trials)
    if args.straggler_group >= 0:
        ctx.set_option(abi.OPT_STRAGGLER_GROUP, args.straggler_group)
    stream = torch.cuda.current_stream()

    def step():
        ctx.fit_device(d_frames.data_ptr(), w.nx, w.ny, w.nframes, d_cands.data_ptr(), cands.size,
                       d_out.data_ptr(), stream.cuda_stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for k in range(args.steps):
        starts[k].record(stream)
        step()
        ends[k].record(stream)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    kernel_ms = float(np.mean([s.elapsed_time(e) for s, e in zip(starts, ends)]))

    fits = d_out.cpu().numpy().view(abi.FIT_DTYPE)
    accepted = int(np.sum(fits["status"] == 0))
    ms_per_step = elapsed * 1e3 / args.steps
    value = accepted * args.steps / elapsed
    flops = algorithmic_flops(fits, cands, w.half_width[0])
    nbytes = algorithmic_bytes(fits, cands, w.half_width[0])
    rbytes = record_bytes(fits, cands, w.half_width[0])
    achieved = flops / (kernel_ms * 1e-3) / 1e12
    print(f"accepted {accepted}/{cands.size}, {ms_per_step:.3f} ms/step, event {kernel_ms:.3f} ms, "
          f"{achieved:.2f} TFLOP/s algorithmic, mean trials {fits['trials'].mean():.2f}",
          file=sys.stderr, flush=True)

    pcie = pcie_inclusive(ctx, frames, cands)
    print(f"pcie-inclusive (host buffers): {pcie['ms_per_call']:.2f} ms per call, {pcie['value']:.4g} loc/s",
          file=sys.stderr, flush=True)
    controller = None
    if not args.no_controller:
        controller = controller_leg(ctx, w, frames, cands, fits, lambda m: print(m, file=sys.stderr, flush=True))
    parity = None
    if not args.no_cpu:
        parity = parity_vs_oracle(w, frames, cands, fits, nthreads)
        print(f"parity vs FP64 oracle over {cands.size} fits: status {parity['status_agreement']:.6f}, x/y within "
              f"1e-3 px {parity['xy_frac_within_1e-3']:.6f} (max {parity['xy_max_px']:.2e} px)",
              file=sys.stderr, flush=True)
    render = None
    if not args.no_render:
        render = render_leg(d_out, cands.size, w, lambda m: print(m, file=sys.stderr, flush=True))
    del d_frames, d_cands, d_out
    torch.cuda.empty_cache()
    mixed = None
    if not args.no_mixed:
        mixed = mixed_leg(max(1, min(args.steps, 5)), 1, lambda m: print(m, file=sys.stderr, flush=True),
                          frames_per_block=args.block_frames, parity=not args.no_cpu)
        torch.cuda.empty_cache()
    dense = None
    if not args.no_dense:
        dense = dense_leg(args.dense_frames, args.dense_block, 100, lambda m: print(m, file=sys.stderr, flush=True),
                          pipelines=args.dense_pipelines)
        torch.cuda.empty_cache()
    net = None
    if not args.no_net:
        net = net_init_leg(args.net_train_rois, args.net_eval_rois, args.net_epochs, max(1, min(args.steps, 5)),
                           1, lambda m: print(m, file=sys.stderr, flush=True))

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(w, frames, cands, args.cpu_seconds, nthreads)
        print(f"cpu baseline {cpu['value']:.0f} loc/s ({cpu['sample']})", file=sys.stderr, flush=True)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "localizations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": w.name,
            "frames": w.nframes,
            "frame_px": f"{w.nx}x{w.ny}",
            "rois": int(cands.size),
            "window": f"{2 * w.half_width[0] + 1}x{2 * w.half_width[0] + 1}",
            "model": "symmetric 5-parameter LM",
            "initial_psf_width": w.initial_psf_width,
            "method": "plain LM (centroid start)",
            "accepted": accepted,
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_TFLOPS,
            "traffic": MEASURED_TRAFFIC_BYTES.get(w.name),
            "kernel_ms": kernel_ms,
            "algorithmic_flops_per_launch": flops,
            "algorithmic_bytes_per_launch": nbytes,
            "algorithmic_bytes_per_roi": nbytes / cands.size,
            "record_bytes_per_launch": rbytes,
            "gather_gbs": nbytes / (kernel_ms * 1e-3) / 1e9,
            "launch": "one lzr_fit_batch_device call (queue sort, fitting kernels, straggler kernels, epilogue)",
        },
        "cpu_baseline": cpu,
        "parity_vs_fp64_oracle": parity,
        "pcie_inclusive": pcie,
        "controller_pattern": controller,
        "net_init": net,
        "mixed": mixed,
        "render": render,
        "dense": dense,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
