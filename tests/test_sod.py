"""The reference's inviscid shock-tube test (ouxsbli/tests/test_st.py: Sod's
problem of 2D_solver/ST, SLAU ORDER 6 HR-SLAU2, hybrid TVD, Euler, TVD-RK3,
density and pressure L1 errors below 1 % of the exact solution) on the 3-D
periodic path, as a quasi-one-dimensional double Riemann problem with the
reference's spacing, time step and end time (tests/golden/make_sod.py).

This pins the SLAU flux, the hybrid switch and the limited MUSCL together
against an exact solution, on the oracle (CPU) and on the GPU library.  The
exact Riemann solver is itself pinned against Toro's tabulated star state of
his Test 1 (Riemann Solvers and Numerical Methods for Fluid Dynamics, 3rd ed.,
Table 4.3: p* = 0.30313, u* = 0.92745, rho*L = 0.42632, rho*R = 0.26557).
"""
import pathlib
import sys

import numpy as np
import pytest

from ouxsbli_amd.cases import CaseSpec, periodic_fill
from oracle_lib import Oracle

GOLD = pathlib.Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(GOLD))
import make_sod  # noqa: E402

L1_TOL = 0.01


def test_exact_solver_against_toro_test1():
    ps, us = make_sod.star_state(1.0, 0.0, 1.0, 0.125, 0.0, 0.1)
    assert ps == pytest.approx(0.30313, abs=5e-6)
    assert us == pytest.approx(0.92745, abs=5e-6)
    rl = make_sod.sample(0.5 * us, 1.0, 0.0, 1.0, 0.125, 0.0, 0.1)[0]
    rr = make_sod.sample(1.1 * us, 1.0, 0.0, 1.0, 0.125, 0.0, 0.1)[0]
    assert rl == pytest.approx(0.42632, abs=5e-6)
    assert rr == pytest.approx(0.26557, abs=5e-6)
    # undisturbed states outside the fan and the shock, a mirror symmetry
    assert make_sod.sample(-2.0, 1.0, 0.0, 1.0, 0.125, 0.0, 0.1) == (1.0, 0.0, 1.0)
    assert make_sod.sample(2.0, 1.0, 0.0, 1.0, 0.125, 0.0, 0.1) == (0.125, 0.0, 0.1)
    a = make_sod.sample(-0.3, 1.0, 0.0, 1.0, 0.125, 0.0, 0.1)
    b = make_sod.sample(0.3, 0.125, 0.0, 0.1, 1.0, 0.0, 1.0)
    assert a[0] == pytest.approx(b[0]) and a[1] == pytest.approx(-b[1]) and a[2] == pytest.approx(b[2])


def test_fixture_is_the_exact_solution():
    g = np.load(GOLD / "sod_double.npz")
    P = make_sod.st_parameters()
    rho, u, p = make_sod.double_problem(g["x"], P["endT"], P)
    np.testing.assert_array_equal(rho, g["rho"])
    np.testing.assert_array_equal(p, g["p"])


def sod_spec(n_transverse, precision="fp64"):
    P = make_sod.st_parameters()
    n = P["n"]
    dx = 2.0 * P["Lx"] / n
    order = 6
    return CaseSpec(nx=n + order, ny=n_transverse + order, nz=n_transverse + order, order=order, scheme="SLAU",
                    visc="Euler", visc_order=0, tvd="hybrid", slau_variant="HRSLAU2", rk=3, precision=precision,
                    gamma=1.4, R=P["R"], threshold=0.4, dt=P["dt"], Lx=2.0 * P["Lx"],
                    Ly=n_transverse * dx, Lz=n_transverse * dx), P


def sod_state(spec, P):
    g = np.load(GOLD / "sod_double.npz")
    x = g["x"]
    high = (x > 0.5 * P["Lx"]) & (x < 1.5 * P["Lx"])
    rho = np.where(high, P["rho0"], P["rho1"])
    p = np.where(high, P["p0"], P["p1"])
    h = spec.order // 2
    Q = np.zeros(spec.state_shape())
    Q[:, :, 0, h:-h] = rho
    Q[:, :, 4, h:-h] = p / (spec.gamma - 1.0)
    return periodic_fill(Q, spec.order)


def l1_errors(Q, spec):
    """Relative L1 errors of density and pressure on each half of the domain
    (each holds one Riemann problem), along the middle line."""
    g = np.load(GOLD / "sod_double.npz")
    h = spec.order // 2
    line = Q[spec.nz // 2, spec.ny // 2, :, h:-h]
    rho = line[0]
    u = line[1] / rho
    p = (spec.gamma - 1.0) * (line[4] - 0.5 * rho * u * u)
    n = rho.size
    out = []
    for half in (slice(0, n // 2), slice(n // 2, n)):
        out.append((np.mean(np.abs(rho[half] - g["rho"][half])) / np.mean(g["rho"][half]),
                    np.mean(np.abs(p[half] - g["p"][half])) / np.mean(g["p"][half])))
    return out


def test_sod_oracle_l1():
    spec, P = sod_spec(8)
    o = Oracle(spec)
    o.set_state(sod_state(spec, P))
    o.step(P["nt"])
    Q = o.get_state()
    # quasi-one-dimensional: every transverse line is the same
    h = spec.order // 2
    np.testing.assert_allclose(Q[h:-h, h:-h], np.broadcast_to(Q[h:h + 1, h:h + 1], Q[h:-h, h:-h].shape),
                               rtol=1e-12, atol=1e-9)
    for l1_rho, l1_p in l1_errors(Q, spec):
        assert l1_rho < L1_TOL and l1_p < L1_TOL, (l1_rho, l1_p)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32_bulk"])
def test_sod_gpu_l1(precision):
    from ouxsbli_amd.solver import Solver
    spec, P = sod_spec(12, precision)
    Q0 = sod_state(spec, P)
    with Solver(spec) as s:
        s.set_state(Q0)
        s.step(P["nt"])
        Q = s.get_state()
    for l1_rho, l1_p in l1_errors(Q, spec):
        assert l1_rho < L1_TOL and l1_p < L1_TOL, (precision, l1_rho, l1_p)


@pytest.mark.gpu
def test_sod_gpu_fp64_equals_oracle_early_steps():
    from ouxsbli_amd.solver import Solver
    spec, P = sod_spec(12)
    Q0 = sod_state(spec, P)
    o = Oracle(spec)
    o.set_state(Q0)
    o.step(40)
    with Solver(spec) as s:
        s.set_state(Q0)
        s.step(40)
        np.testing.assert_array_equal(s.get_qj(), o.get_qj())
