"""CPU checks of the C ABI library: it loads and exports every declared symbol."""
import ctypes
import os
import re

import porousmediaanalysis_amd as pma

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "pma", "pma.h")).read()
    return sorted(set(re.findall(r"^\s*(?:pma_status|void|const char\*)\s+(pma_\w+)\s*\(", text, re.M)))


def test_header_matches_exports_list():
    assert declared() == sorted(pma.EXPORTS)


def test_library_exports_every_symbol():
    lib = ctypes.CDLL(pma.LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), name
