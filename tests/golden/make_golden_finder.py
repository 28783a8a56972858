"""Writes tests/golden/ref_finder.npz: particles found by the reference's own
ParticleFinder_adjacent8::findPositions (source/ParticleFinding.cpp:283-427, built from
the reference sources by oracle/ref/Makefile into oracle/_ref/ref_finder) on synthetic
frames and masks.  Run in the build container (the reference is not on the GPU box):

    make -C oracle/ref && python tests/golden/make_golden_finder.py
"""
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from localizer_amd import synth  # noqa: E402
from oracle import pyoracle as po  # noqa: E402

REF_FINDER = os.path.join(ROOT, "oracle", "_ref", "ref_finder")


def cases():
    rng = np.random.default_rng(1234)
    out = []
    # dense synthetic frames segmented by the oracle GLRT
    w = synth.DENSE_2048
    from dataclasses import replace
    small = replace(w, nx=96, ny=80, nframes=3, emitters_per_frame=25)
    frames, _, _ = synth.generate(small)
    for f in frames:
        mask, _ = po.glrt_mask(f, 1.3, 20.0)
        out.append((f, mask))
    # hand-made masks: diagonal chains, frame edges, single pixels, a ring, a full row
    f = rng.integers(80, 140, (40, 52)).astype(np.uint16)
    m = np.zeros_like(f, np.uint8)
    for k in range(8):
        m[5 + k, 3 + k] = 1           # 8-connected diagonal
    m[0, :6] = 1                      # touches the top edge
    m[10:15, 0] = 1                   # touches the left edge
    m[39, 51] = 1                     # corner pixel
    m[20:25, 20:25] = 1
    m[21:24, 21:24] = 0               # ring (hole inside)
    m[30, 10:50] = 1                  # long row
    m[33, 30] = m[34, 31] = m[35, 30] = 1  # zig-zag
    f[m == 1] += rng.integers(0, 500, int(m.sum())).astype(np.uint16)
    out.append((f, m))
    # random sparse masks
    for _ in range(3):
        f = rng.integers(0, 3000, (33, 47)).astype(np.uint16)
        m = (rng.random((33, 47)) < 0.15).astype(np.uint8)
        out.append((f, m))
    return out


def main():
    cs = cases()
    with tempfile.TemporaryDirectory() as d:
        results = []
        for f, m in cs:
            fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
            ny, nx = f.shape
            with open(fin, "wb") as fh:
                fh.write(struct.pack("<iii", 1, nx, ny))
                fh.write(f.astype("<u2").tobytes())
                fh.write(m.astype(np.uint8).tobytes())
            subprocess.run([REF_FINDER, fin, fout], check=True)
            raw = open(fout, "rb").read()
            (c,) = struct.unpack_from("<i", raw, 0)
            results.append(np.frombuffer(raw, "<f8", 4 * c, 4).reshape(c, 4))
    np.savez_compressed(
        os.path.join(ROOT, "tests", "golden", "ref_finder.npz"),
        n=np.array([len(cs)]),
        **{f"frame{k}": f for k, (f, _) in enumerate(cs)},
        **{f"mask{k}": m for k, (_, m) in enumerate(cs)},
        **{f"particles{k}": r for k, r in enumerate(results)},
    )
    print("cases", len(cs), "particles", [len(r) for r in results])


if __name__ == "__main__":
    main()
