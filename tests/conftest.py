import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

SMALL_CASES = ["small_s1", "small_s2", "small_s3"]
SHIPPED_CASES = ["case_simple-ring", "case_cube-sphere", "case_complex", "case_addax-and-meerkat"]
FULL_CASES = SMALL_CASES + SHIPPED_CASES + ["cube_n8"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (runs through libsbgpu.so)")
    config.addinivalue_line("markers", "slow: minutes of CPU time")


def load_golden(name: str) -> dict:
    path = os.path.join(GOLDEN, name + ".npz")
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def golden_exists(name: str) -> bool:
    return os.path.exists(os.path.join(GOLDEN, name + ".npz"))


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle

    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def gpu():
    """The GPU context. On a GPU run a missing library or device is an error,
    never a skip."""
    from solidboolean_amd import native

    native.load()
    n = native.device_count()
    if n < 1:
        pytest.fail("no HIP device visible to libsbgpu.so")
    ctx = native.Context(0)
    yield ctx
    ctx.close()
