"""Writes tests/golden/ref_storage.npz: the reference's own positions containers
(source/Storage.cpp:268-345 symmetric, :405-476 fixed width, :614-700 ellipsoidal, built by
oracle/ref/Makefile into oracle/_ref/ref_storage) on crafted records -- fitted values, integers,
six-significant-digit rounding cases, large and tiny magnitudes, negative, zero, inf and NaN
fields, frame numbers up to 10^9 -- and an empty container.  Stored per case: the input records,
the matrix of getLocalizedPositionsAsMatrix and the bytes of writePositionsToFile.
Run in the build container (needs the reference sources):
    make -C oracle/ref && python tests/golden/make_golden_storage.py"""
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REF = os.path.join(ROOT, "oracle", "_ref", "ref_storage")
NDOUBLES = {0: 10, 1: 8, 5: 14}
HEADER = "X SIZE:256\nY SIZE:256\nFitting method:Levenberg-Marquardt"


def cases():
    rng = np.random.default_rng(20261204)
    out = []
    special = [0.0, -0.0, 1.0, -1.5, 123456.789, 1234567.0, 0.000123456789, 1e-7, 1e12, 3.0e-300,
               float("inf"), float("-inf"), float("nan"), 99999.95, 0.1 + 0.2]
    for t, nd in NDOUBLES.items():
        n = 40
        frames = rng.integers(0, 10000, n).astype(np.uint64)
        frames[:3] = [0, 1, 10 ** 9]
        nfr = np.ones(n, np.uint64)
        nfr[5] = 7
        vals = np.empty((n, nd))
        vals[:, 0] = rng.uniform(100, 50000, n)          # integral
        vals[:, 1:] = rng.uniform(0.01, 255.0, (n, nd - 1))
        for k, v in enumerate(special):                   # one special value per record and column
            vals[10 + k, k % nd] = v
        out.append((t, frames, nfr, vals))
        out.append((t, frames[:0], nfr[:0], vals[:0]))    # empty: "No positions localized!"
    return out


def write_input(path, t, frames, nfr, vals):
    with open(path, "wb") as f:
        f.write(struct.pack("<ii", t, len(frames)))
        for k in range(len(frames)):
            f.write(struct.pack("<QQ", int(frames[k]), int(nfr[k])))
            f.write(vals[k].astype("<f8").tobytes())


def run(exe, t, frames, nfr, vals):
    with tempfile.TemporaryDirectory() as d:
        fin, fmat, ftxt = (os.path.join(d, x) for x in ("in.bin", "m.bin", "t.txt"))
        write_input(fin, t, frames, nfr, vals)
        r = subprocess.run([exe, fin, fmat, ftxt, HEADER], capture_output=True, timeout=60)
        mat = np.fromfile(fmat, np.uint8) if os.path.exists(fmat) else np.zeros(0, np.uint8)
        txt = np.fromfile(ftxt, np.uint8) if os.path.exists(ftxt) else np.zeros(0, np.uint8)
        return r.returncode, mat, txt, r.stderr.decode()


def main():
    if not os.path.exists(REF):
        sys.exit(f"{REF} missing: make -C oracle/ref")
    arrays = {"header": np.frombuffer(HEADER.encode(), np.uint8)}
    for i, (t, frames, nfr, vals) in enumerate(cases()):
        rc, mat, txt, err = run(REF, t, frames, nfr, vals)
        arrays[f"c{i}_type"] = np.array(t)
        arrays[f"c{i}_frames"] = frames
        arrays[f"c{i}_nframes"] = nfr
        arrays[f"c{i}_values"] = vals
        arrays[f"c{i}_rc"] = np.array(rc)
        arrays[f"c{i}_matrix"] = mat
        arrays[f"c{i}_text"] = txt
        arrays[f"c{i}_stderr"] = np.frombuffer(err.encode(), np.uint8)
        print(f"case {i}: type {t}, {len(frames)} records, rc {rc}, text {txt.size} B {err.strip()}")
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "ref_storage.npz"), **arrays)


if __name__ == "__main__":
    main()
