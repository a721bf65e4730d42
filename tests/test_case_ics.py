"""The ETGV and KHI initial states of the Case layer (ouxsbli_amd/ic.py)
against a scalar restatement of the reference's set_init loops, cell by
cell, with math-library sin/cos:

- ETGV: 3D_solver/ETGV/set.f90:30-40 (rho0, rho0 M0 sin x cos y cos z, ...,
  p = 1/gamma + rho0 M0^2/16 (cos 2x + cos 2y)(cos 2z + 2));
- KHI: 3D_solver/KHI/set.f90:31-47 (stream 1 where y > 0.75 Lx or
  y < 0.25 Lx: the reference compares y with Lx, not Ly, and the grids below
  have Lx != Ly so that the other reading fails; v from x, w from z).

Cell centres follow set_grid_cyclic (src/set_coordinate.f90:120-339), and the
ghost layers must be the periodic image (set_bc_cyclic).  No GPU.
"""
import math

import numpy as np
import pytest

from ouxsbli_amd import ic
from ouxsbli_amd.cases import CaseSpec, cell_centres


def _spec(n, order, L):
    return CaseSpec(nx=n[0] + order, ny=n[1] + order, nz=n[2] + order, order=order, Lx=L[0], Ly=L[1], Lz=L[2],
                    gamma=1.4)


def _etgv_cell(x, y, z, rho0, M0, gam):
    q1 = rho0
    q2 = rho0 * M0 * math.sin(x) * math.cos(y) * math.cos(z)
    q3 = -rho0 * M0 * math.cos(x) * math.sin(y) * math.cos(z)
    q4 = 0.0
    q5 = (1.0 / gam + 0.0625 * rho0 * (M0 ** 2) * (math.cos(2.0 * x) + math.cos(2.0 * y)) * (math.cos(2.0 * z) + 2.0)) \
        / (gam - 1.0) + 0.5 * (q2 ** 2 + q3 ** 2 + q4 ** 2) / q1
    return (q1, q2, q3, q4, q5)


def _khi_cell(x, y, z, Lx, Lz, g, gam):
    pi = math.acos(-1.0)
    w = g["amp"] * g["u1"] * math.sin(2.0 * pi * z / Lz)
    v = g["amp"] * g["u1"] * math.sin(2.0 * pi * x / Lx)
    if y > 0.75 * Lx or y < 0.25 * Lx:
        rho, u = g["rho1"], g["u1"]
    else:
        rho, u = g["rho2"], g["u2"]
    return (rho, rho * u, rho * v, rho * w, g["p"] / (gam - 1.0) + 0.5 * rho * (u ** 2 + v ** 2 + w ** 2))


def _check(spec, Q, cell_fn):
    h = spec.order // 2
    x = cell_centres(spec.nx, spec.Lx, spec.order)
    y = cell_centres(spec.ny, spec.Ly, spec.order)
    z = cell_centres(spec.nz, spec.Lz, spec.order)
    ix, iy, iz = spec.nx - 2 * h, spec.ny - 2 * h, spec.nz - 2 * h
    rng = np.random.default_rng(3)
    picks = [(h, h, h), (spec.nx - h - 1, spec.ny - h - 1, spec.nz - h - 1)]
    picks += [tuple(int(v) for v in (rng.integers(h, h + ix), rng.integers(h, h + iy), rng.integers(h, h + iz)))
              for _ in range(40)]
    # every y row once (the band edges included)
    picks += [(h + 1, j, h + 2) for j in range(h, h + iy)]
    for (i, j, k) in picks:
        want = cell_fn(x[i], y[j], z[k])
        got = Q[k, j, :, i]
        np.testing.assert_allclose(got, want, rtol=4e-16, atol=1e-15 * max(abs(v) for v in want),
                                   err_msg=f"cell {(i, j, k)}")
    # ghost layers: periodic image of the interior
    def img(a, n):
        return a + n if a < h else (a - n if a >= h + n else a)
    for (i, j, k) in [(0, 0, 0), (spec.nx - 1, h, h + 1), (h, spec.ny - 1, 0), (1, h + 3, spec.nz - 1)]:
        np.testing.assert_array_equal(Q[k, j, :, i], Q[img(k, iz), img(j, iy), :, img(i, ix)])


@pytest.mark.parametrize("order", [4, 6])
def test_etgv_state_matches_set_init(order):
    spec = _spec((12, 10, 14), order, (2 * math.pi, 2 * math.pi, 2 * math.pi))
    g = {"rho0": 1.0, "m0": 0.1}
    Q = ic.etgv_state(spec, g)
    _check(spec, Q, lambda x, y, z: _etgv_cell(x, y, z, g["rho0"], g["m0"], spec.gamma))


@pytest.mark.parametrize("order", [4, 6])
def test_khi_state_matches_set_init(order):
    # Ly = 1.6 Lx: the stream-2 band is 0.25..0.75 Lx in y, which a reading
    # with Ly would move to 0.4..1.2
    spec = _spec((16, 26, 12), order, (1.0, 1.6, 0.8))
    g = {"u1": 0.5, "rho1": 1.0, "u2": -0.5, "rho2": 2.0, "p": 2.5, "amp": 0.01}
    Q = ic.khi_state(spec, g)
    _check(spec, Q, lambda x, y, z: _khi_cell(x, y, z, spec.Lx, spec.Lz, g, spec.gamma))
    # both streams present, split where the reference's Lx test puts it
    y = cell_centres(spec.ny, spec.Ly, spec.order)
    h = order // 2
    inner = (y[h:-h] > 0.25 * spec.Lx) & (y[h:-h] < 0.75 * spec.Lx)
    rho_col = Q[h + 1, h:-h, 0, h + 1]
    np.testing.assert_array_equal(rho_col, np.where(inner, 2.0, 1.0))


def test_device_ic_routing():
    """NSTGV, ETGV and KHI build their set_init on the device (ic.device_ic);
    an explicit ic= and the other cases stay on the host."""
    from ouxsbli_amd import _abi
    spec = _spec((12, 10, 14), 4, (2 * math.pi, 2 * math.pi, 2 * math.pi))
    d = ic.device_ic("ETGV", spec, {"rho0": 1.2, "m0": 0.1})
    assert (d.kind, d.rho0, d.V0) == (_abi.OX_IC_ETGV, 1.2, 0.1)
    g = {"u1": 0.5, "rho1": 1.0, "u2": -0.5, "rho2": 2.0, "p": 2.5, "amp": 0.01}
    d = ic.device_ic("KHI", spec, g)
    assert (d.kind, d.V0, d.rho0, d.u2, d.rho2, d.p0, d.amp) == (_abi.OX_IC_KHI, 0.5, 1.0, -0.5, 2.0, 2.5, 0.01)
    assert ic.device_ic("KHI", spec, {**g, "ic_type": 1}) is None
    assert ic.device_ic("STZ", spec, {}) is None


def test_ox_ic_struct_layout():
    """ox_ic: two int32, a uint64, seven doubles, a pointer, three doubles."""
    import ctypes
    from ouxsbli_amd import _abi
    assert ctypes.sizeof(_abi.OxIc) == 4 + 4 + 8 + 7 * 8 + 8 + 3 * 8
    assert _abi.OxIc.u2.offset == 4 + 4 + 8 + 7 * 8 + 8
