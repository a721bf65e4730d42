"""The HIP library against the committed golden fixtures of the oracle
(tests/golden/): the FP64 right-hand side and 3-step state of each case must
hash to the oracle's digests, i.e. be bit-identical, without running the
oracle."""
import json
import pathlib
import sys

import numpy as np
import pytest

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(GOLDEN))

import make_golden as mg  # noqa: E402
from ouxsbli_amd.solver import Solver  # noqa: E402

pytestmark = pytest.mark.gpu

META = json.loads((GOLDEN / "oracle_golden.json").read_text())
BY_ID = {e["case"]: e for e in META["cases"]}


@pytest.mark.parametrize("c", mg.CASES, ids=mg.case_id)
def test_gpu_matches_golden_digests(c):
    spec = mg.spec_of(c)
    Q = mg.input_state(spec)
    with Solver(spec) as s:
        s.set_state(Q)
        R = mg.interior(spec, s.rhs())
        s.step(META["steps"])
        S = s.get_qj()
    e = BY_ID[mg.case_id(c)]
    assert mg.digest(R) == e["rhs_sha256"], "rhs differs from the golden digest"
    assert mg.digest(S) == e["state_sha256"], "state differs from the golden digest"
