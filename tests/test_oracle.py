"""The oracle (oracle/sb_oracle.c) against the reference's golden vectors,
brute-force closed forms and hand-worked known answers (CPU only)."""
import pytest

import oracle_checks as OC
from conftest import FULL_CASES

CASE_CHECKS = ["normals", "boxes", "tree", "pairs_dfs_order", "gd_status", "gd_segments", "rays", "classify"]


@pytest.fixture(scope="module")
def case_results(oracle_lib):
    return {name: OC.check_case(oracle_lib.lib(), name) for name in FULL_CASES}


@pytest.mark.parametrize("name", FULL_CASES)
@pytest.mark.parametrize("check", CASE_CHECKS)
def test_case(case_results, name, check):
    assert case_results[name][check], f"oracle differs from the reference on {name}: {check}"


def test_gd_vectors(oracle_lib):
    assert OC.check_gd_vectors(oracle_lib.lib())["gd_vectors"]


@pytest.mark.parametrize("check", ["brute_force_pairs", "sphere_inside", "gd_known_answer", "gd_parallel_coplanar",
                                   "position_key"])
def test_closed_form(oracle_lib, check):
    assert OC.check_closed_forms(oracle_lib.lib())[check]
