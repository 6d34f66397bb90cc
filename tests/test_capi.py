"""The C ABI library loads and exports every entry point include/gomc_hip.h
declares (no device calls)."""
import ctypes

from gomc_amd import _lib


def test_library_exports_header_symbols():
    lib = _lib.load()
    names = _lib.header_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


def test_no_device_means_error_code_not_crash():
    # Creating a box with an invalid descriptor is rejected on the host.
    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.gomc_box_create(None, 0, ctypes.byref(h)) < 0
