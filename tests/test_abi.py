"""The C-ABI library loads and exports every function include/tblite_amd.h
declares; the Python binding covers the same set.  No GPU calls."""
import ctypes
import os
import re

import pytest

import tblite_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "tblite_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tblite_amd_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(tblite_amd.LIBNAME)
    missing = [name for name in _declared() if not hasattr(lib, name)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(_declared()) == set(tblite_amd.EXPORTED)


def test_default_parameters_through_abi():
    p = tblite_amd.default_param()
    assert p.nelem == 5
    assert [p.elem[i].z for i in range(5)] == [1, 6, 7, 8, 16]
    assert abs(p.kshell[0][0] - 1.85) < 1e-15


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(OSError):
        tblite_amd.load(str(tmp_path / "libtblite_amd.so"))
