"""Builds the native parts in-tree.

    python -m localizer_amd.build          # everything
products:
    localizer_amd/lib/liblocalizer_hip.so  HIP kernels + C ABI (gfx950)
    localizer_amd/lib/liblzr_synth.so      synthetic workload generator (host C)
    tests/emu/liblzr_emu.so                host build of the lane logic (tests only)
    oracle/liblzr_oracle.so                FP64 oracle (tests / CPU baseline only)
    oracle/_ref/ref_models                 reference model functions (when the
                                           reference sources are present)
    tests/_san/                            sanitizer builds of the host code
                                           (`python -m localizer_amd.build sanitize`)
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
REFERENCE = "/root/reference"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
OFFLOAD_ARCH = "gfx950"
# Packed FP32 instructions stay off: every VALU op is a plain (non-packed) FP32/FP64 op.
HIP_FLAGS = [
    "-O3",
    "-std=c++17",
    f"--offload-arch={OFFLOAD_ARCH}",
    "-fPIC",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_synth():
    os.makedirs(LIB, exist_ok=True)
    _run(["gcc", "-O2", "-fPIC", "-std=c11", "-Wall", "-shared", "-o",
          os.path.join(LIB, "liblzr_synth.so"), os.path.join(CSRC, "synth.c"), "-lm", "-lpthread"])


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def build_ref():
    if os.path.isdir(os.path.join(REFERENCE, "source")):
        _run(["make", "-s", "-C", os.path.join(ROOT, "oracle", "ref"), f"REF={REFERENCE}"])


def build_emu():
    """Host build of lzr_lane.h for the CPU emulation used by the tests."""
    emu = os.path.join(ROOT, "tests", "emu")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
          "-o", os.path.join(emu, "liblzr_emu.so"), os.path.join(emu, "lzr_emu.cpp")])


def build_hip():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "liblocalizer_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, os.path.join(CSRC, "lzr_fit.hip"), os.path.join(CSRC, "lzr_net.hip"),
          os.path.join(CSRC, "lzr_find.hip"), os.path.join(CSRC, "lzr_pipeline.hip"),
          os.path.join(CSRC, "lzr_render.hip"), os.path.join(CSRC, "lzr_batch.hip")])
    os.replace(tmp, out)


def build_dropin_driver():
    """C++ test driver of the drop-in classes (links the HIP library)."""
    cpp = os.path.join(ROOT, "tests", "cpp")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-o", os.path.join(cpp, "dropin_main"),
          os.path.join(cpp, "dropin_main.cpp"), f"-L{LIB}", "-llocalizer_hip",
          "-Wl,-rpath,$ORIGIN/../../localizer_amd/lib", "-lpthread"])
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-o", os.path.join(cpp, "storage_main"),
          os.path.join(cpp, "storage_main.cpp")])
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-o", os.path.join(cpp, "render_main"),
          os.path.join(cpp, "render_main.cpp"), f"-L{LIB}", "-llocalizer_hip",
          "-Wl,-rpath,$ORIGIN/../../localizer_amd/lib", "-lpthread"])


SAN = os.path.join(ROOT, "tests", "_san")
ASAN = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"]
TSAN = ["-O1", "-g", "-fsanitize=thread"]


def build_sanitized():
    """Host code under the sanitizers (tests only), into tests/_san/: the oracle and the lane
    emulation as shared libraries and the drop-in positions containers' driver with
    AddressSanitizer + UndefinedBehaviorSanitizer, and a ThreadSanitizer driver of the
    oracle's frame-parallel pool (tests/cpp/tsan_main.cpp)."""
    os.makedirs(SAN, exist_ok=True)
    # each product is linked to a temporary name and renamed into place, so a test process that
    # runs or maps the previous build keeps a whole file
    def out(name):
        return os.path.join(SAN, name + ".tmp")

    def publish(name):
        os.replace(out(name), os.path.join(SAN, name))

    orc = [os.path.join(ROOT, "oracle", f) for f in ("lzr_oracle.c", "lzr_find_oracle.c", "lzr_render_oracle.c")]
    _run(["gcc", *ASAN, "-fPIC", "-std=c11", "-shared", "-o", out("liblzr_oracle.so"), *orc, "-lm", "-lpthread"])
    publish("liblzr_oracle.so")
    emu = os.path.join(ROOT, "tests", "emu")
    _run(["g++", *ASAN, "-std=c++17", "-fPIC", "-shared", "-Wno-unknown-pragmas", "-o", out("liblzr_emu.so"),
          os.path.join(emu, "lzr_emu.cpp")])
    publish("liblzr_emu.so")
    cpp = os.path.join(ROOT, "tests", "cpp")
    _run(["g++", *ASAN, "-std=c++17", "-o", out("storage_main"), os.path.join(cpp, "storage_main.cpp")])
    publish("storage_main")
    objs = []
    for src in orc:
        obj = os.path.join(SAN, os.path.basename(src)[:-2] + "_tsan.o")
        _run(["gcc", *TSAN, "-std=c11", "-c", "-o", obj, src])
        objs.append(obj)
    _run(["g++", *TSAN, "-std=c++17", "-o", out("tsan_main"), os.path.join(cpp, "tsan_main.cpp"), *objs,
          "-lm", "-lpthread"])
    publish("tsan_main")


def build_host_libs():
    build_synth()
    build_oracle()
    build_emu()


def build_all():
    build_host_libs()
    build_ref()
    build_hip()
    build_dropin_driver()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "hip": build_hip, "host": build_host_libs, "ref": build_ref,
         "emu": build_emu, "oracle": build_oracle, "synth": build_synth, "dropin": build_dropin_driver,
         "sanitize": build_sanitized}[w]()
