"""C ABI checks that need no GPU: the library loads, exports every symbol the
header declares, and reports DeviceNotFound / validation errors the way the
reference's NttError does."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sfntt.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sfn_[a-z_0-9]+)\s*\(", text)))


def test_header_parses_and_is_nonempty():
    syms = declared_symbols()
    assert "sfn_ntt" in syms and "sfn_coset_lde_batch" in syms and len(syms) >= 25


def test_library_exports_every_declared_symbol():
    from sfntt import _lib
    lib = _lib.load()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    # the Python binding declares a signature for every header symbol
    assert set(declared_symbols()) == set(_lib.SIGNATURES)


def test_status_strings():
    from sfntt import _lib
    lib = _lib.load()
    assert lib.sfn_status_string(0) == b"ok"
    assert lib.sfn_status_string(1) == b"invalid size"
    assert lib.sfn_status_string(2) == b"device not found"


def test_ctx_create_reports_device_not_found_without_gpu():
    from sfntt import _lib
    lib = _lib.load()
    p = ctypes.c_void_p()
    rc = lib.sfn_ctx_create(0, ctypes.byref(p))
    if rc == 0:  # a GPU is present: the context must be usable and destroyable
        lib.sfn_ctx_destroy(p)
    else:
        assert rc == _lib.ERR_DEVICE_NOT_FOUND
        assert not p.value
    assert lib.sfn_ctx_create(10_000, ctypes.byref(p)) == _lib.ERR_DEVICE_NOT_FOUND
    assert lib.sfn_ctx_create(0, None) == _lib.ERR_INVALID_ARGUMENT


def test_null_context_is_rejected():
    from sfntt import _lib
    lib = _lib.load()
    assert lib.sfn_ntt(None, 0, 0, 0, 4, 1, 0, None, None, None) == _lib.ERR_INVALID_ARGUMENT
    assert lib.sfn_coset_lde_batch(None, 4, 1, 1, 0, 0, None, None, None) == _lib.ERR_INVALID_ARGUMENT
    assert lib.sfn_last_error(None) == b""
    assert lib.sfn_call_timing(None, 1) == _lib.ERR_INVALID_ARGUMENT
    ns = ctypes.c_uint64()
    assert lib.sfn_last_device_ns(None, ctypes.byref(ns)) == _lib.ERR_INVALID_ARGUMENT
    assert lib.sfn_pass_timing(None, 1) == _lib.ERR_INVALID_ARGUMENT


def test_python_api_raises_device_not_found_or_works():
    import sfntt
    try:
        ctx = sfntt.HipContext(0)
    except sfntt.api.DeviceNotFound:
        return
    assert ctx.device_info()["gcn_arch"].startswith("gfx950")
