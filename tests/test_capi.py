"""The C-ABI library loads on a host without a GPU and exports every symbol
declared in include/libatssim_hip.h; calls that need no device work."""
import ctypes
import os
import re

import numpy as np

from libatssim_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "libatssim_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lats_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_exported():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    assert set(declared_symbols()) == set(_lib.SIGNATURES)


def test_errors_are_reported():
    L = _lib.lib()
    rc = L.lats_vbd_create(None, None, None)
    assert rc == _lib.LATS_ERR_INVALID_ARGUMENT
    assert b"null" in L.lats_last_error()


def test_create_without_device_fails_loudly():
    """No silent CPU path: creating a context on a host without a GPU fails."""
    import pytest
    hip = ctypes.CDLL("libamdhip64.so")
    n = ctypes.c_int(0)
    if hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0:
        pytest.skip("a GPU is present")
    from libatssim_amd import configs, GpuVBD
    cloth, params, _, _ = configs.grid_c0()
    with pytest.raises(_lib.LatsError) as ei:
        GpuVBD(cloth, params)
    assert ei.value.code == _lib.LATS_ERR_NO_DEVICE
