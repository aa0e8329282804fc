"""The C ABI library loads and exports every function include/sbgpu.h declares
(no device calls: this runs without a GPU)."""
import ctypes
import os
import re

from solidboolean_amd import native

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "sbgpu.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(sbg_\w+)\s*\(", text, re.M)))


def test_header_and_binding_agree():
    assert declared() == sorted(native.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(native.LIB_PATH)
    missing = [s for s in declared() if not hasattr(lib, s)]
    assert not missing, missing
