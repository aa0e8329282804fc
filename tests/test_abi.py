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


def test_mesh_index_range_checked_before_the_library():
    """Indices of 32-bit and floating types that do not fit the ABI's uint32
    are refused in Python, before any library call (the stand-in context has
    no library); 64-bit integer indices go to sbg_mesh_create64, which checks
    them itself."""
    import types

    import numpy as np
    import pytest

    ctx = types.SimpleNamespace(L=None, h=None)
    v = np.zeros((4, 3))
    for dtype in (np.int32, np.float64):
        for bad in (-1, 2 ** 32, 2 ** 40):
            if np.dtype(dtype).itemsize == 4 and bad >= 2 ** 31:
                continue
            t = np.array([[0, 1, 2], [1, 2, 3], [0, 2, bad]], dtype)
            with pytest.raises(ValueError, match="uint32 range"):
                native.Mesh(ctx, v, t)
            with pytest.raises(ValueError, match="uint32 range"):
                native.Mesh(ctx, v, t[:, ::-1])  # strided view
