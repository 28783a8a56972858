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
    xi = (4, 5) if model == abi.MODEL_ELLIPSOIDAL else (3, 4)
    pi = (3, 4) if model == abi.MODEL_ELLIPSOIDAL else (2, 3)
    got = {(int(r[0]), round(r[xi[0]], 2), round(r[xi[1]], 2)) for r in m}
    hits = 0
    for k, r in ref:
        key = (k, round(r["params"][pi[0]], 2), round(r["params"][pi[1]], 2))
        hits += key in got
    _say(f"pipeline: {hits} of {len(ref)} oracle rows found")
    assert hits >= 0.98 * len(ref)
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
