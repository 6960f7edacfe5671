"""The C-ABI library loads on a machine without a GPU and exports every declared symbol;
context creation reports an error code (no device) instead of crashing."""
import ctypes
import os

import pytest

from mpsdyn_amd import device as dev


def test_library_built_and_loads():
    assert os.path.exists(dev.LIB_PATH), "run python -m mpsdyn_amd.build"
    dev.load()


def test_exports_every_declared_symbol():
    lib = dev.load()
    declared = dev._declared_symbols()
    assert len(declared) >= 20
    missing = [n for n in declared if not hasattr(lib, n)]
    assert not missing, f"declared in include/*.h but not exported: {missing}"


def test_null_arguments_are_errors():
    lib = dev.load()
    assert lib.mpsd_ctx_create(0, None) == -1
    assert lib.mpsd_ctx_destroy(None) == -1
    assert lib.mpsd_tensor_destroy(None) == -1
    assert lib.mpsd_heff_apply(None, None, None, 1, 1, 1, 0) == -1
    assert lib.mpsd_last_error(None) == b"null context"


def test_ctx_create_without_device_fails_cleanly():
    if os.environ.get("HIP_VISIBLE_DEVICES") is None and _has_gpu():
        pytest.skip("a GPU is present")
    lib = dev.load()
    h = ctypes.c_void_p()
    assert lib.mpsd_ctx_create(0, ctypes.byref(h)) < 0


def _has_gpu():
    return os.path.exists("/dev/kfd")
