"""Writes tests/golden/ref_models.npz: golden model residuals, Jacobians and 2x2
eigen-decompositions produced by the reference's own code.

The residual/Jacobian values come from FitFunction_SymmetricGaussian,
Jacobian_SymmetricGaussian, FitFunction_EllipsoidalGaussian and
Jacobian_EllipsoidalGaussian (source/Localization.cpp:1007-1295) compiled from
/root/reference by oracle/ref/Makefile into oracle/_ref/ref_models; the eigen values
come from the vendored Eigen 3.3.7 SelfAdjointEigenSolver the reference calls at
source/Localization.cpp:525-547.  Run from the repository root:

    make -C oracle/ref && python tests/golden/make_golden.py
"""
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_models")
OUT = os.path.join(ROOT, "tests", "golden", "ref_models.npz")


def make_cases(rng):
    cases = []
    for kind in (0, 1):
        for h in (2, 3, 4, 5, 7):
            n = 2 * h + 1
            for c in range(5):
                xoff = float(rng.integers(0, 2000))
                yoff = float(rng.integers(0, 2000))
                sigma = 1.0 if c % 2 == 0 else 2.5
                th = np.zeros(7)
                if kind == 0:
                    th[:5] = [rng.uniform(50, 2000), rng.uniform(0.8, 2.5),
                              xoff + h + rng.uniform(-1.5, 1.5), yoff + h + rng.uniform(-1.5, 1.5),
                              rng.uniform(20, 200)]
                else:
                    th[:] = [rng.uniform(50, 2000), rng.uniform(0.8, 2.5), rng.uniform(0.8, 2.5),
                             xoff + h + rng.uniform(-1.5, 1.5), yoff + h + rng.uniform(-1.5, 1.5),
                             rng.uniform(-0.6, 0.6), rng.uniform(20, 200)]
                    if c == 4:
                        th[2] = th[1]  # sigma_x == sigma_y: the rho column is exact there
                data = rng.integers(0, 3000, size=n * n).astype(np.float64)
                cases.append((kind, n, xoff, yoff, sigma, th, data))
        # zero width: the reference returns GSL_FAILURE
        th = np.zeros(7)
        th[:5] = [100.0, 0.0, 3.0, 3.0, 10.0] if kind == 0 else [100.0, 0.0, 1.0, 3.0, 3.0, 0.0, 10.0][:5]
        if kind == 1:
            th[:] = [100.0, 0.0, 1.0, 3.0, 3.0, 0.0, 10.0]
        cases.append((kind, 7, 0.0, 0.0, 1.0, th, np.zeros(49)))
    eig = [(1.3, 1.3, 0.0), (1.0, 2.0, 0.0), (2.0, 1.0, 0.0), (1.3, 1.3, 0.3), (1.3, 1.3, -0.3)]
    for _ in range(40):
        eig.append((rng.uniform(0.5, 3.0), rng.uniform(0.5, 3.0), rng.uniform(-0.95, 0.95)))
    for sx, sy, rho in eig:
        th = np.zeros(7)
        th[:3] = [sx, sy, rho]
        cases.append((2, 1, 0.0, 0.0, 1.0, th, np.zeros(1)))
    return cases


def run_harness(cases):
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("<i", len(cases)))
            for kind, n, xoff, yoff, sigma, th, data in cases:
                f.write(struct.pack("<ii", kind, n))
                f.write(struct.pack("<ddd", xoff, yoff, sigma))
                f.write(np.asarray(th, "<f8").tobytes())
                f.write(np.asarray(data, "<f8").tobytes())
        subprocess.run([HARNESS, fin, fout], check=True)
        raw = open(fout, "rb").read()
    res, pos = [], 0
    for kind, n, *_ in cases:
        sf, sj = struct.unpack_from("<ii", raw, pos)
        pos += 8
        if kind == 2:
            vals = np.frombuffer(raw, "<f8", 3, pos).copy()
            pos += 24
            res.append((sf, sj, vals, None))
            continue
        N, p = n * n, 5 if kind == 0 else 7
        f = np.frombuffer(raw, "<f8", N, pos).copy()
        pos += 8 * N
        J = np.frombuffer(raw, "<f8", N * p, pos).copy().reshape(N, p)
        pos += 8 * N * p
        res.append((sf, sj, f, J))
    return res


def main():
    if not os.path.exists(HARNESS):
        sys.exit(f"{HARNESS} missing: make -C oracle/ref")
    rng = np.random.default_rng(20240603)
    cases = make_cases(rng)
    res = run_harness(cases)
    out = {}
    mi = [i for i, c in enumerate(cases) if c[0] != 2]
    ei = [i for i, c in enumerate(cases) if c[0] == 2]
    out["kind"] = np.array([cases[i][0] for i in mi], np.int32)
    out["n"] = np.array([cases[i][1] for i in mi], np.int32)
    out["xoff"] = np.array([cases[i][2] for i in mi])
    out["yoff"] = np.array([cases[i][3] for i in mi])
    out["sigma"] = np.array([cases[i][4] for i in mi])
    out["theta"] = np.stack([cases[i][5] for i in mi])
    out["status_f"] = np.array([res[i][0] for i in mi], np.int32)
    out["status_J"] = np.array([res[i][1] for i in mi], np.int32)
    data = np.concatenate([cases[i][6] for i in mi])
    fv = np.concatenate([res[i][2] for i in mi])
    Jv = np.concatenate([res[i][3].reshape(-1) for i in mi])
    out["data"], out["f"], out["J"] = data, fv, Jv
    out["eig_in"] = np.stack([cases[i][5][:3] for i in ei])
    out["eig_info"] = np.array([res[i][0] for i in ei], np.int32)
    out["eig_out"] = np.stack([res[i][2] for i in ei])
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT}: {len(mi)} model cases, {len(ei)} eigen cases")


if __name__ == "__main__":
    main()
