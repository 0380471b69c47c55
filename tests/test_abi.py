"""The C-ABI library loads and exports every entry point include/fc3.h declares.

CPU-only: no compute calls.  Entry points that need a device are exercised only
for their argument validation, which happens before any HIP call.
"""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fc3.h")


def header_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fc3_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from fc3 import _lib
    L = _lib.load()
    names = header_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(L, name), name
    # the ctypes signature table covers exactly the header
    assert sorted(_lib.SIGNATURES) == names


def test_version_and_status_strings():
    from fc3 import _lib
    L = _lib.load()
    assert L.fc3_version() >= 100
    assert L.fc3_status_string(0) == b"ok"
    assert L.fc3_status_string(3) == b"invalid block_size"
    assert L.fc3_status_string(99) == b"unknown status"


def test_null_handle_is_rejected():
    from fc3 import _lib
    L = _lib.load()
    st = L.fc3_synchronize(None)
    assert st == _lib.FC3_ERR_ARGUMENT
    assert b"null" in L.fc3_last_error()
    st = L.fc3_mcc3(None, None, None, 8, 0, None)
    assert st == _lib.FC3_ERR_ARGUMENT


def test_check_maps_assertions():
    from fc3 import _lib
    L = _lib.load()
    L.fc3_synchronize(None)
    with pytest.raises(AssertionError):
        _lib.check(_lib.FC3_ERR_ARGUMENT)
    with pytest.raises(_lib.FC3Error):
        _lib.check(_lib.FC3_ERR_HIP)


def test_library_is_gfx950_code_object():
    # the shared object carries an amdgcn gfx950 code object
    from fc3 import _lib
    data = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in data
