"""GPU candidate generation (GLRT + 8-way components) against the oracle
(oracle/lzr_find_oracle.c, pinned by the reference's own particle finder)."""
import sys
from dataclasses import replace

import numpy as np
import pytest

from localizer_amd import abi, device, native, synth
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _oracle(frames, width, pfa):
    out, masks = [], []
    for k, f in enumerate(frames):
        m, _ = po.glrt_mask(f, width, pfa)
        p = po.find_particles(f, m)
        masks.append(m)
        out.append(np.column_stack([p, np.full(len(p), k)]))
    return np.concatenate(out) if out else np.zeros((0, 5)), masks


@pytest.mark.parametrize("nx,ny,width,pfa", [(256, 192, 1.3, 20.0), (255, 201, 1.0, 25.0), (160, 128, 2.0, 15.0)])
def test_gpu_candidates_match_oracle(nx, ny, width, pfa):
    _say(f"case: finder {nx}x{ny} w={width} PFA={pfa}, 6 dense frames")
    w = replace(synth.DENSE_2048, nx=nx, ny=ny, nframes=6, emitters_per_frame=60)
    frames, _, _ = synth.generate(w)
    ref, _ = _oracle(frames, width, pfa)
    d_frames = device.upload(frames)
    cap = 20000
    d_cands = device.empty(abi.CANDIDATE_DTYPE, cap)
    import torch

    d_count = torch.zeros(1, dtype=torch.int32, device="cuda")
    with native.Finder(width, pfa) as fd:
        fd.find_device(d_frames.ptr, nx, ny, len(frames), d_cands.ptr, cap, d_count.data_ptr(), device.stream_handle())
        torch.cuda.synchronize()
        n = int(d_count.item())
    got = d_cands.numpy()[:n]
    _say(f"finder: gpu {n} candidates, oracle {len(ref)}")
    assert n == len(ref)
    np.testing.assert_array_equal(got["frame"], ref[:, 4].astype(np.int32))
    np.testing.assert_array_equal(got["x"], ref[:, 0])
    np.testing.assert_array_equal(got["y"], ref[:, 1])
    np.testing.assert_array_equal(got["intensity"], ref[:, 2])
    np.testing.assert_array_equal(got["background"], ref[:, 3])
    assert np.all(got["half_width"] == 0) and np.all(got["model"] == -1)


def test_gpu_finder_overflow_and_empty():
    import torch

    _say("case: finder capacity overflow and an empty frame stack")
    rng = np.random.default_rng(2)
    frames = rng.integers(0, 60000, (2, 96, 96)).astype(np.uint16)  # noise: many segmented pixels
    d_frames = device.upload(frames)
    d_cands = device.empty(abi.CANDIDATE_DTYPE, 8)
    d_count = torch.zeros(1, dtype=torch.int32, device="cuda")
    with native.Finder(1.3, 1.0, max_segmented_per_frame=64) as fd:
        fd.find_device(d_frames.ptr, 96, 96, 2, d_cands.ptr, 8, d_count.data_ptr(), device.stream_handle())
        torch.cuda.synchronize()
        assert int(d_count.item()) == -1
        fd.find_device(d_frames.ptr, 96, 96, 0, d_cands.ptr, 8, d_count.data_ptr(), device.stream_handle())
        torch.cuda.synchronize()
        assert int(d_count.item()) == 0
