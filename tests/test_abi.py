"""The C-ABI library loads and exports exactly what include/emerge_hip.h declares."""
import re
from pathlib import Path

import pytest

from emerge_amd import _lib

HEADER = Path(__file__).resolve().parents[1] / "include" / "emerge_hip.h"


def declared_functions() -> set[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(emg_[a-z0-9_]+)\s*\(", text))


def test_header_and_binding_agree():
    assert declared_functions() == set(_lib.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert lib.emg_version().decode().count(".") == 2


def test_create_without_device_fails_loudly():
    import ctypes
    lib = _lib.load()
    h = ctypes.c_void_p()
    # on a machine without a HIP device this must be an error, never a CPU path
    rc = lib.emg_create(10_000, ctypes.byref(h))
    assert rc != 0


def test_null_arguments_rejected():
    lib = _lib.load()
    assert lib.emg_create(0, None) != 0
    assert lib.emg_last_error(None, None, 0) != 0
