"""The C ABI library loads and exports every entry point include/mars_hip.h declares (no GPU calls)."""

import ctypes

import pytest

from tests import abi


def test_header_declares_entry_points():
    syms = abi.declared_symbols()
    for s in ("mars_hip_context_create", "mars_hip_orthonormalize", "mars_hip_orthonormalize_batch",
              "mars_hip_fit_create", "mars_hip_fit_eval", "mars_hip_fit_append"):
        assert s in syms


@pytest.mark.parametrize("sym", abi.declared_symbols())
def test_library_exports(sym):
    lib = ctypes.CDLL(abi.LIB)
    assert hasattr(lib, sym), sym


def test_error_path_without_device_call():
    lib = abi.load()
    # argument validation happens before any device work
    assert lib.mars_hip_context_destroy(None) == 0
    rc = lib.mars_hip_fit_info(None, None, None, None)
    assert rc == -1
    assert b"null context" in lib.mars_hip_last_error()
