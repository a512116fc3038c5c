"""The C ABI library loads and exports every symbol include/tdcc.h declares (no GPU calls)."""
import ctypes
import os
import re

import pytest

from python_code_to_check_c_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "tdcc.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tdcc_[a-z0-9_]+)\s*\(", text)))


def test_header_matches_binding_list():
    assert declared_symbols() == sorted(_lib.EXPORTS)


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="libtdcc_hip.so not built")
def test_library_exports_all_declared_symbols():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name
