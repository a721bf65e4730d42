"""The serving workload (Kelvin-Helmholtz requests) against the oracle, and
the device reductions that size and guard a solve (CFL wave speed, the
non-physical-state scan).

- A Kelvin-Helmholtz request (Euler, RK3; KEEP ORDER 4 and SLAU ORDER 6 with
  HR-SLAU2, the scheme families the stream draws) admitted into the batch
  engine with its initial state built on the device, as the server admits
  it, ends bit-identical to the FP64 oracle started from that same state.
- Served FP32-bulk results: relative kinetic-energy error against the FP64
  oracle below 1e-5 (the bar of the FP32-bulk form).
- max(|u|,|v|,|w|) + c over the interior: exactly the oracle's value and a
  numpy maximum over the same expression, and NaN (not a finite number a
  caller would size dt from) when a cell has negative pressure.
- ox_check_state: OX_ESTATE with the first bad linear cell index and the
  count, for a negative density and a NaN energy planted in the state.
"""
import math

import numpy as np
import pytest

from ouxsbli_amd import cases, serve
from ouxsbli_amd.batch import Batch
from ouxsbli_amd.solver import Solver
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

GIB = 1 << 30


def _device_ic_qj(spec):
    """The QJ the library builds for the request on the device."""
    with Solver(spec) as s:
        s.init_state()
        return s.get_qj()


@pytest.mark.parametrize("order,scheme", [(4, "KEEP"), (6, "SLAU"), (4, "SLAU"), (6, "KEEP")])
def test_khi_request_fp64_equals_oracle(order, scheme):
    spec = cases.khi_spec(32, order, scheme, "fp64", seed=21)
    nsteps = 6
    QJ0 = _device_ic_qj(spec)
    o = Oracle(spec)
    o.set_qj(QJ0)
    o.step(nsteps)
    with Batch(hbm_budget=4 * GIB, max_cases=4, autotune=False) as b:
        a = b.admit(spec, None, nsteps)
        other = cases.khi_spec(24, 10 - order, "KEEP" if scheme == "SLAU" else "SLAU", "fp64", seed=22)
        b.admit(other, None, nsteps + 2)
        while a not in b.done():
            b.step()
        Qg = b.retire(a)
    np.testing.assert_array_equal(Qg, o.get_state())


def test_served_fp32_ke_error_below_1e5():
    reqs = serve.khi_stream(5, rate_hz=100.0, seed=31, n_min=16, n_max=48, nsteps=20, precision="fp32_bulk",
                            with_states=False)
    rep = serve.serve(reqs, hbm_budget=4 * GIB, max_cases=3, autotune=False)
    worst = 0.0
    for r in rep.requests:
        assert r.error is None and r.result is not None
        spec64 = cases.khi_spec(r.spec.nx - r.spec.order, r.spec.order, r.spec.scheme, "fp64", seed=r.spec.seed)
        o = Oracle(spec64)
        o.set_qj(_device_ic_qj(spec64))
        o.step(r.nsteps)
        ke_ref = o.diagnostics()["ke"]
        og = Oracle(spec64)
        og.set_state(r.result)
        err = abs(og.diagnostics()["ke"] - ke_ref) / abs(ke_ref)
        worst = max(worst, err)
        assert err < 1e-5, (r.rid, r.spec.scheme, r.spec.order, err)
    assert worst > 0.0  # FP32 flux arithmetic does differ from FP64


def _jac(spec):
    dx = spec.Lx / (spec.nx - spec.order)
    dy = spec.Ly / (spec.ny - spec.order)
    dz = spec.Lz / (spec.nz - spec.order)
    return 8.0 / ((dx + dx) * (dy + dy) * 2.0 * dz)


def _wave_speed_numpy(spec, QJ):
    h = spec.order // 2
    q = QJ[h:-h, h:-h, :, h:-h]
    J = _jac(spec)
    q1 = q[:, :, 0, :]
    rho = J * q1
    u, v, w = q[:, :, 1, :] / q1, q[:, :, 2, :] / q1, q[:, :, 3, :] / q1
    p = (spec.gamma - 1.0) * (J * q[:, :, 4, :] - 0.5 * rho * (u * u + v * v + w * w))
    cs = np.sqrt(spec.gamma * p / rho)
    return float(np.max(np.maximum(np.abs(u), np.maximum(np.abs(v), np.abs(w))) + cs))


@pytest.mark.parametrize("order", [2, 6])
def test_max_wave_speed_exact(order):
    spec = cases.tgv_spec(18, order, Ma=0.9, cfl=0.3, cfl_mode="acoustic", scheme="SLAU", visc="Euler", rk=3)
    spec.ny, spec.nz = 14 + order, 22 + order
    Q = np.ascontiguousarray(cases.periodic_fill(_random_state(spec, seed=order), order))
    with Solver(spec) as s:
        s.set_state(Q)
        QJ = s.get_qj()
        dg = s.diagnostics()
    o = Oracle(spec)
    o.set_state(Q)
    do = o.diagnostics()
    ref = _wave_speed_numpy(spec, QJ)
    assert dg["max_wave_speed"] == ref
    assert do["max_wave_speed"] == ref
    # the maximum is attained where it should be: a lone fast cell (u = 10,
    # the same pressure) moves it
    h = order // 2
    Q2 = Q.copy()
    c = (h + 5, h + 3, slice(None), h + 7)
    rho = Q2[c][0]
    u, v, w = Q2[c][1] / rho, Q2[c][2] / rho, Q2[c][3] / rho
    p = (spec.gamma - 1.0) * (Q2[c][4] - 0.5 * rho * (u * u + v * v + w * w))
    Q2[h + 5, h + 3, 1, h + 7] = rho * 10.0
    Q2[h + 5, h + 3, 4, h + 7] = p / (spec.gamma - 1.0) + 0.5 * rho * (100.0 + v * v + w * w)
    with Solver(spec) as s:
        s.set_state(Q2)
        assert s.diagnostics()["max_wave_speed"] == _wave_speed_numpy(spec, s.get_qj()) > ref


def _random_state(spec, seed):
    rng = np.random.default_rng(seed)
    shape = (spec.nz, spec.ny, spec.nx)
    rho = 1.0 + 0.2 * rng.random(shape)
    u, v, w = (rng.standard_normal(shape) * 0.3 for _ in range(3))
    p = 1.0 + 0.5 * rng.random(shape)
    Q = np.empty(spec.state_shape())
    Q[:, :, 0, :] = rho
    Q[:, :, 1, :] = rho * u
    Q[:, :, 2, :] = rho * v
    Q[:, :, 3, :] = rho * w
    Q[:, :, 4, :] = p / (spec.gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
    return Q


def test_nonphysical_state_is_reported():
    spec = cases.khi_spec(16, 4, "KEEP", "fp64", seed=3)
    Q = cases.khi_state(spec)
    h = spec.order // 2
    with Solver(spec) as s:
        s.set_state(Q)
        assert s.check_state() == (True, -1, 0)
        assert math.isfinite(s.diagnostics()["max_wave_speed"])
        bad = Q.copy()
        bad[h + 4, h + 2, 0, h + 9] = -1.0                    # negative density
        bad[h + 7, h + 1, 4, h + 3] = float("nan")            # NaN energy
        bad[h + 2, h + 6, 4, h + 5] = 0.0                     # negative pressure (kinetic > total energy)
        s.set_state(bad)
        ok, idx, cnt = s.check_state()
        assert not ok and cnt == 3
        lin = lambda i, j, k: i + spec.nx * (j + spec.ny * k)
        # numpy index [k, j, l, i]; the library's linear index is x fastest
        assert idx == min(lin(h + 9, h + 2, h + 4), lin(h + 3, h + 1, h + 7), lin(h + 5, h + 6, h + 2))
        assert math.isnan(s.diagnostics()["max_wave_speed"])
        # the library's status for the same scan
        from ouxsbli_amd import _abi
        import ctypes
        i64 = ctypes.c_int64()
        rc = s._lib.ox_check_state(s._h, ctypes.byref(i64), ctypes.byref(ctypes.c_int64()))
        assert rc == _abi.OX_ESTATE and i64.value == idx
