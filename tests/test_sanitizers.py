"""The host code under the sanitizers: the oracle, the lane emulation and the drop-in
positions containers built with AddressSanitizer + UndefinedBehaviorSanitizer run the
CPU tests that exercise them; the oracle's frame-parallel pool (the reference controller's
one-frame-per-task pattern, source/PALMAnalysis.cpp:204-252) and concurrent per-frame
finder calls run under ThreadSanitizer.  Builds: `python -m localizer_amd.build sanitize`
(tests/_san/, rebuilt here when a source is newer)."""
import fcntl
import glob
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "tests", "_san")
PRODUCTS = ["liblzr_oracle.so", "liblzr_emu.so", "storage_main", "tsan_main"]
SOURCES = (glob.glob(os.path.join(ROOT, "oracle", "*.[ch]")) + glob.glob(os.path.join(ROOT, "tests", "emu", "*.cpp"))
           + glob.glob(os.path.join(ROOT, "tests", "cpp", "*.cpp")) + glob.glob(os.path.join(ROOT, "oracle", "ref", "*.h"))
           + glob.glob(os.path.join(ROOT, "localizer_amd", "csrc", "lzr_lane.h"))
           + glob.glob(os.path.join(ROOT, "localizer_amd", "dropin", "StorageHIP.h")))
ASAN_FILES = ["test_emu_parity.py", "test_finder_oracle.py", "test_oracle_golden.py", "test_oracle_minpack.py",
              "test_oracle_record.py", "test_render_oracle.py", "test_sep_pass.py", "test_storage_golden.py",
              "test_lm_convergence.py"]


def _runtime(name):
    p = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


@pytest.fixture(scope="module")
def built():
    if not (_runtime("libasan.so") and _runtime("libtsan.so")):
        pytest.skip("gcc sanitizer runtimes not installed")
    newest = max(os.path.getmtime(f) for f in SOURCES)
    os.makedirs(SAN, exist_ok=True)
    # one process rebuilds at a time (parallel test workers); the others wait, then see it fresh
    with open(os.path.join(SAN, ".build.lock"), "w") as lock:
        fcntl.flock(lock, fcntl.LOCK_EX)
        stale = any(not os.path.exists(os.path.join(SAN, p)) or os.path.getmtime(os.path.join(SAN, p)) < newest
                    for p in PRODUCTS)
        if stale:
            subprocess.run([sys.executable, "-m", "localizer_amd.build", "sanitize"], cwd=ROOT, check=True,
                           capture_output=True, timeout=600)
    return SAN


def _asan_env():
    env = dict(os.environ)
    env.update(LD_PRELOAD=_runtime("libasan.so"), ASAN_OPTIONS="detect_leaks=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1",
               LZR_ORACLE_LIB=os.path.join(SAN, "liblzr_oracle.so"), LZR_EMU_LIB=os.path.join(SAN, "liblzr_emu.so"),
               LZR_STORAGE_MAIN=os.path.join(SAN, "storage_main"))
    return env


def test_sanitized_libraries_are_the_ones_loaded(built):
    code = ("from oracle import pyoracle as po\nfrom tests.emu import pyemu as pe\npo.lib(); pe.lib()\n"
            "print(open('/proc/self/maps').read())")
    out = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=_asan_env(), capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr[-2000:]
    for lib in ("liblzr_oracle.so", "liblzr_emu.so"):
        assert os.path.join(SAN, lib) in out.stdout
    assert "libasan" in out.stdout


def test_cpu_suite_under_address_and_undefined_sanitizers(built):
    files = [os.path.join(ROOT, "tests", f) for f in ASAN_FILES]
    out = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-p", "no:xdist",
                          "-m", "not gpu", *files], cwd=ROOT, env=_asan_env(), capture_output=True, text=True,
                         timeout=900)
    tail = (out.stdout + out.stderr)[-3000:]
    assert out.returncode == 0, tail
    assert "ERROR: AddressSanitizer" not in out.stderr and "runtime error:" not in out.stderr, tail
    print(out.stdout.strip().splitlines()[-1])


def test_oracle_pool_under_thread_sanitizer(built):
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1")
    out = subprocess.run([os.path.join(SAN, "tsan_main")], env=env, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr[-3000:]
    assert "ThreadSanitizer" not in out.stderr, out.stderr[-3000:]
    assert "ok" in out.stdout
