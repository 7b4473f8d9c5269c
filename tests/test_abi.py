"""The C-ABI library loads and exports every symbol include/strut_gpu.h
declares (no compute calls: this runs without a GPU)."""

import ctypes
import os
import re

from strut_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(sgpu_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_native.LIB_PATH)
    declared = _declared()
    assert declared, "no declarations found"
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_native.EXPORTED) == declared


def test_binding_loads_and_reports_abi_version():
    lib = _native.load()
    assert lib.sgpu_abi_version() == 1


def test_struct_layouts_match_header():
    assert ctypes.sizeof(_native.SectionPoint) == 32
    assert ctypes.sizeof(_native.SectionResult) == 8 * 15 + 8
