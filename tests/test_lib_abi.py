"""The C ABI library loads and exports exactly what include/tdtm.h declares (CPU only)."""

import ctypes
import os
import re

import pytest

from leopard_em_amd import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "tdtm.h")


def _declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(tdtm_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_bound_symbols():
    assert _declared() == sorted(_lib.SIGNATURES)


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="libtdtm.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in _declared():
        assert hasattr(lib, name), name


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="libtdtm.so not built")
def test_plan_size_validation_errors_before_any_device_call():
    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.tdtm_match_plan_create(4000, 4096, 512, 8, 13, 1, ctypes.byref(h)) != 0
    assert b"power of two" in lib.tdtm_last_error()
    assert lib.tdtm_match_plan_create(256, 256, 512, 8, 1, 1, ctypes.byref(h)) != 0
    assert b"template" in lib.tdtm_last_error()
    assert lib.tdtm_match_plan_create(256, 256, 64, 0, 1, 1, ctypes.byref(h)) != 0
    assert lib.tdtm_abi_version() == 1
    with pytest.raises(_lib.TdtmError):
        _lib.check(lib.tdtm_match_plan_create(256, 256, 48, 1, 1, 1, ctypes.byref(h)))


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(_lib.TdtmError, match="not built"):
        _lib.load(str(tmp_path / "nope.so"))
