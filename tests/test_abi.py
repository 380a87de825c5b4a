"""The C-ABI library loads and exports every function include/usbs_hip.h
declares (no GPU needed: nothing is called)."""
import ctypes
import os
import re

from usbs_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "usbs_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(usbs_\w+)\s*\(", text, flags=re.M)))


def test_header_symbols_exported():
    names = declared_functions()
    assert "usbs_solve" in names and "usbs_lanczos_smallest" in names
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_version_string():
    assert _lib.lib().usbs_version().startswith(b"usbs-hip")
