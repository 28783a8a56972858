"""The C-ABI library loads and exports every function include/localizer_hip.h declares;
without a GPU, lzr_create fails with a clear error instead of falling back."""
import ctypes
import os
import re

import pytest

from localizer_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            names += re.findall(r"LZR_EXPORT\s+\w+\s*\**\s*(lzr_\w+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_the_binding_set():
    assert declared_functions() == sorted(native.EXPORTED)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(native.LIB_PATH):
        pytest.skip("HIP library not built")
    L = ctypes.CDLL(native.LIB_PATH)
    for name in declared_functions():
        assert hasattr(L, name), name
    assert native.lib().lzr_abi_version() == 1


def test_create_without_device_raises():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(native.LocalizerError, match="no HIP device"):
        native.Context(0, 1.3)


def test_invalid_constructor_arguments_raise():
    with pytest.raises(native.LocalizerError):
        native.Context(0, -1.0)
    with pytest.raises(native.LocalizerError):
        native.Context(9, 1.3)
