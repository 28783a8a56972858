"""The drop-in positions containers (localizer_amd/dropin/StorageHIP.h, through
tests/cpp/storage_main) against the reference's own containers (source/Storage.cpp, golden
outputs in tests/golden/ref_storage.npz from tests/golden/make_golden_storage.py): the
matrix of getLocalizedPositionsAsMatrix (12 / 10 / 16 columns) and the bytes of
writePositionsToFile (header lines, column titles, default stream formatting of every
field, inf/nan, frame numbers), for the symmetric, fixed-width and ellipsoidal types, and
the error an empty container raises."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.environ.get("LZR_STORAGE_MAIN") or os.path.join(ROOT, "tests", "cpp", "storage_main")
GOLDEN = np.load(os.path.join(ROOT, "tests", "golden", "ref_storage.npz"), allow_pickle=False)

import sys  # noqa: E402

sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
import make_golden_storage as mk  # noqa: E402

CASES = sorted({int(k.split("_")[0][1:]) for k in GOLDEN.files if k.startswith("c")})


@pytest.mark.parametrize("case", CASES)
def test_dropin_containers_match_reference_bytes(case):
    if not os.path.exists(DRIVER):
        pytest.skip("storage driver not built (localizer_amd.build dropin)")
    g = {k.split("_", 1)[1]: GOLDEN[k] for k in GOLDEN.files if k.startswith(f"c{case}_")}
    t = int(g["type"])
    rc, mat, txt, err = mk.run(DRIVER, t, g["frames"], g["nframes"], g["values"])
    assert rc == int(g["rc"])
    if rc != 0:
        assert err.strip() == bytes(g["stderr"]).decode().strip()
        return
    assert np.array_equal(mat, g["matrix"])
    assert bytes(txt) == bytes(g["text"])
    # the matrix in the reference's column order (source/Storage.cpp:268-290, :405-420, :614-640)
    raw = mat.view(np.uint8)
    rows, cols = np.frombuffer(raw[:16].tobytes(), np.int64)
    m = np.frombuffer(raw[16:].tobytes(), np.float64).reshape(cols, rows).T
    assert cols == {0: 12, 1: 10, 5: 16}[t]
    assert np.array_equal(m[:, 0], g["frames"].astype(np.float64))
    assert np.array_equal(m[:, -1], g["nframes"].astype(np.float64))
