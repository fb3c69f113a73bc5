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


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="libtdtm.so not built")
def test_prep_validation_errors_before_any_device_call():
    lib = _lib.load()
    assert lib.tdtm_prepare_image_workspace_bytes(100, 127, 0) == -1
    assert lib.tdtm_prepare_image_workspace_bytes(4096, 16384, 0) == -1
    assert lib.tdtm_prepare_image_workspace_bytes(64, 64, 1) == -1
    # (3 x 33 bins + 1 filter block) doubles for 64 x 64
    assert lib.tdtm_prepare_image_workspace_bytes(64, 64, 0) == (3 * 33 + 1) * 8
    assert lib.tdtm_prepare_image_workspace_bytes(64, 64, 100) == (3 * 100 + 1) * 8
    fake = ctypes.c_void_p(16)  # never dereferenced: every check below fails first
    assert lib.tdtm_rfft2(None, 64, 64, fake, None) != 0
    assert b"null" in lib.tdtm_last_error()
    assert lib.tdtm_rfft2(fake, 48, 64, fake, None) != 0
    assert b"power of two" in lib.tdtm_last_error()
    assert lib.tdtm_volume_rfft(fake, 1024, fake, None) != 0
    assert b"power of two" in lib.tdtm_last_error()
    assert lib.tdtm_prepare_image(fake, 64, 64, 31, 1, 0, 0.5, fake, fake, fake, None) != 0
    assert b"template size" in lib.tdtm_last_error()
    assert lib.tdtm_prepare_image(fake, 64, 64, 32, 1, 0, 0.0, fake, fake, fake, None) != 0
    assert b"max_freq" in lib.tdtm_last_error()
    params = (ctypes.c_double * 6)(300.0, 2.7, 1.5, 0.0, 1.0, 0.0)  # amplitude contrast >= 1
    assert lib.tdtm_ctf_filter_stack(64, 1, 1, 1e4, 1e4, 0.0, fake, fake, params, fake, None) != 0
    assert b"CTF" in lib.tdtm_last_error()


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="libtdtm.so not built")
def test_regrid_validation_errors_before_any_device_call():
    lib = _lib.load()
    assert lib.tdtm_irfft2_any_workspace_bytes(4092, 5760) == 4092 * 2881 * 16
    assert lib.tdtm_irfft2_any_workspace_bytes(5000, 64) == -1   # height above 4096
    assert lib.tdtm_irfft2_any_workspace_bytes(100, 101) == -1   # odd width
    fake = ctypes.c_void_p(16)  # never dereferenced: every check below fails first
    assert lib.tdtm_irfft2_any(fake, 100, 128, fake, 64, fake, None) != 0
    assert b"stride" in lib.tdtm_last_error()
    assert lib.tdtm_irfft2_any(fake, 100, 8194, fake, 9000, fake, None) != 0
    assert b"irfft2_any" in lib.tdtm_last_error()
    h = ctypes.c_void_p()
    assert lib.tdtm_match_plan_create(4096, 16384, 512, 8, 13, 1, ctypes.byref(h)) != 0
    assert b"power of two" in lib.tdtm_last_error()


def test_search_grid_rule():
    from leopard_em_amd.backend.core_match_template import search_grid

    assert search_grid(4096, 4096) == (4096, 4096)
    assert search_grid(4092, 5760) == (4096, 8192)
    assert search_grid(200, 300) == (256, 512)
    assert search_grid(32, 32) == (64, 64)
    with pytest.raises(_lib.TdtmError, match="exceeds"):
        search_grid(5760, 4092)
