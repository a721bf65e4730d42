"""The oracle against its committed golden fixtures (tests/golden/,
written by tests/golden/make_golden.py): SHA-256 of the FP64 bytes of the
right-hand side and of the 3-step state for eleven scheme / order / viscous /
RK combinations, and the full arrays of two of them."""
import json
import pathlib
import sys

import numpy as np
import pytest

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(GOLDEN))

import make_golden as mg  # noqa: E402
from oracle_lib import Oracle  # noqa: E402

META = json.loads((GOLDEN / "oracle_golden.json").read_text())
ARRAYS = np.load(GOLDEN / "oracle_golden.npz")
BY_ID = {e["case"]: e for e in META["cases"]}


@pytest.mark.parametrize("c", mg.CASES, ids=mg.case_id)
def test_oracle_matches_golden_digests(c):
    spec = mg.spec_of(c)
    Q = mg.input_state(spec)
    o = Oracle(spec)
    o.set_state(Q)
    R = mg.interior(spec, o.rhs())
    o.step(META["steps"])
    S = o.get_qj()
    e = BY_ID[mg.case_id(c)]
    assert mg.digest(R) == e["rhs_sha256"]
    assert mg.digest(S) == e["state_sha256"]
    if c in mg.FULL:
        k = mg.case_id(c)
        np.testing.assert_array_equal(Q, ARRAYS[k + "/input"])
        np.testing.assert_array_equal(R, ARRAYS[k + "/rhs"])
        np.testing.assert_array_equal(S, ARRAYS[k + "/state"])
