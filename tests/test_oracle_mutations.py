"""Each planted mistake in the oracle (ORC_MUTATION in oracle/sb_oracle.c) must
fail at least one of the checks that pin the faithful oracle.  The mutants are
built into a temporary directory, on the CPU."""
import os
import subprocess

import pytest

import oracle_checks as OC
from oracle import oracle as O

SRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "sb_oracle.c")
MUTANTS = {
    1: "open box interval",
    2: "split ties go left",
    3: "transposed endpoint parameter",
    4: ">= in CONSTRUCT_INTERSECTION",
    5: "PositionKey grid 1e-4",
    6: "dropped term in ray-plane dot",
    7: ">= in the inside test",
    8: "vote threshold",
}


@pytest.mark.parametrize("k", sorted(MUTANTS))
def test_mutant_is_caught(tmp_path, k):
    so = str(tmp_path / f"orc_m{k}.so")
    subprocess.check_call(["gcc", "-O2", "-fPIC", "-ffp-contract=off", "-std=c99", "-shared", f"-DORC_MUTATION={k}",
                           "-o", so, SRC, "-lm"])
    res = OC.all_checks(O.load_lib(so))
    failing = [name for name, ok in res.items() if not ok]
    assert failing, f"mutant {k} ({MUTANTS[k]}) passed every check"
