import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


def _ensure_built():
    need = [
        os.path.join(ROOT, "oracle", "liblzr_oracle.so"),
        os.path.join(ROOT, "localizer_amd", "lib", "liblzr_synth.so"),
    ]
    if not all(os.path.exists(p) for p in need):
        from localizer_amd import build

        build.build_host_libs()


_ensure_built()


def pytest_collection_modifyitems(config, items):
    """GPU tests need a HIP device; on a machine without one they are skipped."""
    try:
        import torch

        has_gpu = torch.cuda.is_available()
    except Exception:  # pragma: no cover
        has_gpu = False
    if has_gpu:
        return
    skip = pytest.mark.skip(reason="no HIP device")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def root():
    return ROOT
