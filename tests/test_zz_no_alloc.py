"""ox_diagnostics_noalloc (no device or host allocation per call: its
reduction partials live in the library's static device array) against
ox_diagnostics (partials in the handle's scratch, merged on the host): every
field identical, on two handles in a row, the second with a smaller grid,
each called twice.  The first grid has more interior cells than 4096 blocks
of 256 threads, so both paths run the reduction at its block cap."""
import ctypes
import math

import numpy as np
import pytest

from ouxsbli_amd import _abi, _native, cases
from ouxsbli_amd.solver import Solver

pytestmark = pytest.mark.gpu


def _both(s):
    L = _native.load()
    a, b = _abi.OxDiag(), _abi.OxDiag()
    _native.check(L.ox_diagnostics_noalloc(s._h, ctypes.byref(b)))
    _native.check(L.ox_diagnostics(s._h, ctypes.byref(a)))
    return a.as_dict(), b.as_dict()


def _same(a, b):
    for k in a:
        assert (a[k] == b[k]) or (math.isnan(a[k]) and math.isnan(b[k])), (k, a[k], b[k])


def test_noalloc_diagnostics_equal_allocating_path():
    big = cases.khi_spec(112, 6, "SLAU", "fp64", seed=5)    # 1.4 M interior cells: 4096 blocks
    small = cases.khi_spec(16, 4, "KEEP", "fp64", seed=6)   # 4096 cells: 16 blocks
    assert big.interior_cells > 4096 * 256
    for spec in (big, small):
        with Solver(spec) as s:
            s.init_state()
            s.step(2)
            a, b = _both(s)
            _same(a, b)
            assert a["ke"] > 0 and math.isfinite(a["max_wave_speed"])
            a2, b2 = _both(s)   # the array keeps the last call's contents; a repeat is unaffected
            _same(a2, b2)
            _same(a, a2)


def test_noalloc_diagnostics_nan_state():
    spec = cases.khi_spec(16, 4, "KEEP", "fp64", seed=7)
    Q = cases.khi_state(spec)
    Q[5, 4, 4, 6] = float("nan")
    with Solver(spec) as s:
        s.set_state(Q)
        a, b = _both(s)
    _same(a, b)
    assert math.isnan(b["max_wave_speed"])


def test_noalloc_rejects_null():
    L = _native.load()
    assert L.ox_diagnostics_noalloc(None, None) == _abi.OX_EINVAL
