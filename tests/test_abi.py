"""The C ABI: libphx.so loads without a GPU, exports every function that
include/phx.h declares, and its record layouts match the ctypes mirror."""
import os
import re

from pyheadtail_amd import _phx

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      "include", "phx.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(phx_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _phx.load_library()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_declares_every_symbol():
    names = declared_functions()
    assert sorted(_phx.SIGNATURES) == names


def test_record_layouts_match():
    assert _phx.check_layouts() == {}


def test_host_only_entry_points():
    lib = _phx.load_library()
    assert lib.phx_abi_version() == 2
    assert lib.phx_bunch_sums_len(6) == 51
    assert lib.phx_bunch_sums_len(7) == -1
    assert lib.phx_workspace_bytes(10**6, 500) > 10**6 // 1024 * 64 * 8
    assert lib.phx_engine_ws_bytes(10**6, 500, 2, 1) > 0
    assert lib.phx_sizeof(b"nonexistent") == -1


def test_invalid_arguments_are_reported_without_device():
    lib = _phx.load_library()
    rc = lib.phx_bunch_sums(None, 9, 10, None, None, None)
    assert rc == -1
    assert b"k=9" in lib.phx_last_error()
    rc = lib.phx_transverse_segment(None, None, None, None, None, 10, None, None)
    assert rc == -1
