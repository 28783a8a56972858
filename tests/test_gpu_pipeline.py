"""The whole pipeline (lzr_localize / lzr_fit_emitters: GLRT + 8-way finder + fit) on the
GPU against the oracle pipeline (oracle GLRT, oracle particle finder, FP64 oracle fit)."""
import sys
from dataclasses import replace

import numpy as np
import pytest

from localizer_amd import abi, native, synth
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _oracle_pipeline(frames, w, pfa, model):
    rows = []
    for k, f in enumerate(frames):
        mask, _ = po.glrt_mask(f, w, pfa)
        parts = po.find_particles(f, mask)
        c = np.zeros(len(parts), abi.CANDIDATE_DTYPE)
        c["x"], c["y"], c["intensity"], c["background"] = parts[:, 0], parts[:, 1], parts[:, 2], parts[:, 3]
        c["frame"] = 0
        c["model"] = -1
        fits = po.fit_batch(f[None], c, model, w)
        for r in fits[fits["status"] == 0]:
            rows.append((k, r))
    return rows


def _record_row(k, r, model):
    """The reference matrix row (source/Storage.cpp:268-290 / :614-640) of an oracle fit."""
    p = r["params"]
    d = r["deviations"]
    if model == abi.MODEL_ELLIPSOIDAL:
        return [k, r["integral"], r["sd1"], r["sd2"], p[3], p[4], r["theta"], p[6], r["integral_deviation"],
                r["sd1_deviation"], r["sd2_deviation"], d[3], d[4], 0.0, d[6], 1]
    return [k, r["integral"], p[1], p[2], p[3], p[4], r["integral_deviation"], d[1], d[2], d[3], d[4], 1]


@pytest.mark.parametrize("model", [abi.MODEL_ELLIPSOIDAL, abi.MODEL_SYMMETRIC])
def test_gpu_pipeline_matches_oracle_pipeline(model):
    _say(f"case: pipeline model {model}, 5 frames 192x160, w 1.3, PFA 20")
    wl = replace(synth.DENSE_2048, nx=192, ny=160, nframes=5, emitters_per_frame=40)
    frames, _, _ = synth.generate(wl)
    m = native.localize(frames, 1.3, 20.0, model)
    ref = _oracle_pipeline(frames, 1.3, 20.0, model)
    _say(f"pipeline: gpu {len(m)} rows, oracle {len(ref)} rows")
    assert m.shape[1] == (16 if model == abi.MODEL_ELLIPSOIDAL else 12)
    assert np.all(np.diff(m[:, 0]) >= 0)  # sorted by frame
    assert abs(len(m) - len(ref)) <= max(1, 0.01 * len(ref))
    # the finders are bit-exact, so both pipelines fit the same candidates: pair each oracle row
    # with the GPU row of its frame nearest in position, then compare every column
    xi = (4, 5) if model == abi.MODEL_ELLIPSOIDAL else (3, 4)
    mat = np.array([_record_row(k, r, model) for k, r in ref])
    paired, misses = 0, {}
    for row in mat:
        cand = m[m[:, 0] == row[0]]
        if cand.size == 0:
            continue
        d = np.hypot(cand[:, xi[0]] - row[xi[0]], cand[:, xi[1]] - row[xi[1]])
        g = cand[np.argmin(d)]
        if d.min() > 0.5:
            continue
        paired += 1
        for c in range(1, m.shape[1]):
            if c in xi:
                ok = abs(g[c] - row[c]) <= 1e-3
            elif model == abi.MODEL_ELLIPSOIDAL and c in (6, 13):  # theta, thetaDeviation: absolute
                ok = abs(g[c] - row[c]) <= 1e-4
            else:
                ok = abs(g[c] - row[c]) <= 1e-4 * abs(row[c]) or g[c] == row[c]
            misses[c] = misses.get(c, 0) + (not ok)
    _say(f"pipeline: {paired} of {len(ref)} oracle rows paired; misses per column {misses}")
    assert paired >= len(ref) - 1
    # about 200 rows: at most one miss per column (the 0.1 % bar at scale: tests/test_gpu_parity.py)
    assert max(misses.values()) <= 1, misses
    if model == abi.MODEL_ELLIPSOIDAL:
        n2, cols = None, None
        import ctypes
        n2, p = ctypes.c_int(), ctypes.POINTER(ctypes.c_double)()
        f = np.ascontiguousarray(frames)
        assert native.lib().lzr_fit_emitters(f.ctypes.data, 192, 160, 5, 1.3, 20.0, ctypes.byref(n2),
                                             ctypes.byref(p)) == 0
        flat = np.ctypeslib.as_array(p, shape=(n2.value * 16,)).copy()
        native.lib().lzr_free(p)
        np.testing.assert_array_equal(flat.reshape(16, -1).T, m)
