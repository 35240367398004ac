"""The C ABI library: builds, loads without a GPU and exports every symbol
declared in include/csrst_hip.h."""
import os
import re

from csrst_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "csrst_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(csrst_\w+)\s*\(", text, re.M)))


def test_header_matches_binding_list():
    assert declared() == sorted(_lib.SYMBOLS)


def test_library_exports_all_symbols():
    L = _lib.lib()
    for name in declared():
        assert hasattr(L, name), name
    assert L.csrst_version().startswith(b"csrst")
