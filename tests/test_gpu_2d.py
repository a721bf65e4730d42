"""The 2-D oblique-shock path on the GPU.

The bar: the FP64 form matches the FP64 oracle (oracle/ox_oracle2d.cpp) bit
for bit after many steps, at grids that leave partial tiles, for both
schemes and every order; the FP32-bulk form stays within a stated bound of
the FP64 result; a case advanced in a batch with 2-D and 3-D cases joining
and leaving ends bit-identical to the same case advanced alone; and the
reference's own oblique-shock run (257 x 129, 0.1 ms) reproduces the
Rankine-Hugoniot states of its test (ouxsbli/tests/test_os.py).
"""
import dataclasses
import math

import numpy as np
import pytest

from ouxsbli_amd import cases, cases2d as c2
from ouxsbli_amd.batch import Batch
from ouxsbli_amd.solver import Solver
from ouxsbli_amd.solver2d import Solver2D
from oracle_lib import Oracle2D
from test_oracle2d import os_regions, primitives, reference_reflected_shock

pytestmark = pytest.mark.gpu

GIB = 1 << 30


def noisy_os_state(spec, seed):
    """The OS initial state with seeded 2 % noise in the interior density and
    velocity, so that limiters, the sensor and the wiggle flag all switch."""
    Q = c2.os_state(spec)
    rng = np.random.default_rng(seed)
    sl = (slice(1, -1), slice(1, -1))
    rho = Q[:, 0, :][sl] * (1.0 + 0.02 * rng.uniform(-1, 1, (spec.ny - 2, spec.nx - 2)))
    mu = Q[:, 1, :][sl] / Q[:, 0, :][sl] * (1.0 + 0.02 * rng.uniform(-1, 1, rho.shape))
    mv = 5.0 * rng.uniform(-1, 1, rho.shape)
    p = (c2.GAMMA - 1.0) * (Q[:, 3, :][sl] - 0.5 * Q[:, 1, :][sl] ** 2 / Q[:, 0, :][sl])
    Q[:, 0, :][sl] = rho
    Q[:, 1, :][sl] = rho * mu
    Q[:, 2, :][sl] = rho * mv
    Q[:, 3, :][sl] = p / (c2.GAMMA - 1.0) + 0.5 * rho * (mu * mu + mv * mv)
    return Q


PARITY = [
    # nx, ny, scheme, order, variant, tvd, steps
    (70, 23, "SLAU", 6, "HRSLAU2", "hybrid", 40),
    (131, 37, "Roe", 6, "HRSLAU2", "hybrid", 30),
    (45, 30, "SLAU", 4, "SLAU", "tvd", 30),
    (66, 17, "Roe", 4, "HRSLAU2", "hybrid", 30),
    (40, 21, "SLAU", 2, "HRSLAU2", "none", 30),
    (129, 16, "Roe", 2, "HRSLAU2", "none", 30),
    (64, 9, "SLAU", 6, "HRSLAU2", "none", 20),
]


@pytest.mark.parametrize("nx,ny,scheme,order,variant,tvd,steps", PARITY)
@pytest.mark.parametrize("noisy", [False, True])
def test_fp64_matches_oracle_bitwise(nx, ny, scheme, order, variant, tvd, steps, noisy):
    print(f"2d parity {nx}x{ny} {scheme} ORDER {order} {variant} {tvd} noisy={noisy}", flush=True)
    spec = c2.os_spec(nx, ny, scheme=scheme, order=order, slau_variant=variant, tvd=tvd, cfl=0.3)
    Q0 = noisy_os_state(spec, nx * ny) if noisy else c2.os_state(spec)
    o = Oracle2D(spec)
    o.set_state(Q0)
    o.step(steps)
    ref = o.get_state()
    with Solver2D(spec) as s:
        s.set_state(Q0)
        s.step(steps)
        got = s.get_state()
    assert np.isfinite(got).all()
    assert not np.array_equal(ref, Q0)
    bad = np.argwhere(got != ref)
    assert bad.size == 0, f"{len(bad)} values differ, first at (j, l, i) = {bad[0]}"


@pytest.mark.parametrize("scheme", ["SLAU", "Roe"])
def test_fp32_bulk_stays_close_to_fp64(scheme):
    """After 300 steps of the developing shock reflection on 200 x 90 the
    FP32-bulk fields stay within 3e-4 of the FP64 result in relative L2 norm
    (momentum components against the x-momentum scale), and within 1e-2
    pointwise: at the shock fronts an FP32 rounding can flip a limiter
    switch, which moves single cells by up to ~1e-3 of the field's range."""
    spec = c2.os_spec(200, 90, scheme=scheme, cfl=0.3)
    Q0 = c2.os_state(spec)
    out = {}
    for prec in ("fp64", "fp32_bulk"):
        sp = dataclasses.replace(spec, precision=prec)
        with Solver2D(sp) as s:
            s.set_state(Q0)
            s.step(300)
            out[prec] = s.get_state()
    a, b = out["fp64"], out["fp32_bulk"]
    for l in range(4):
        ref = a[:, 1, :] if l == 2 else a[:, l, :]
        d = a[:, l, :] - b[:, l, :]
        l2 = np.sqrt(np.mean(d * d) / np.mean(ref * ref))
        mx = np.abs(d).max() / np.abs(ref).max()
        print(f"{scheme} var {l}: rel L2 {l2:.3e}  max {mx:.3e}")
        assert l2 < 3e-4
        assert mx < 1e-2


def _alone2d(spec, Q, n):
    with Solver2D(spec) as s:
        s.set_state(Q)
        s.step(n)
        return s.get_state()


def _alone3d(spec, Q, n):
    with Solver(spec) as s:
        s.set_state(Q)
        s.step(n)
        return s.get_state()


@pytest.mark.parametrize("precision", ["fp64", "fp32_bulk"])
@pytest.mark.parametrize("use_graph", [True, False])
def test_2d_alone_equals_in_mix(precision, use_graph):
    """2-D cases of several sizes, schemes and orders, with a 3-D case beside
    them, joining at different steps: each ends bit-identical to its own
    single-case run."""
    mix = [
        (c2.os_spec(150, 70, scheme="Roe", precision=precision, cfl=0.3), 9, 0),
        (c2.os_spec(97, 49, scheme="SLAU", precision=precision, cfl=0.3), 6, 0),
        (c2.os_spec(70, 130, scheme="Roe", order=4, precision=precision, cfl=0.3,
                    beta=c2.weak_beta(2.0, math.radians(11.0))), 5, 2),
        (c2.os_spec(300, 40, scheme="Roe", precision=precision, cfl=0.3,
                    beta=c2.weak_beta(2.0, math.radians(5.0))), 7, 1),
        (cases.khi_spec(24, 4, "SLAU", precision, seed=9), 4, 1),
    ]
    states = [c2.os_state(s) if getattr(s, "dim", 3) == 2 else cases.initial_state(s) for s, _, _ in mix]
    ref = [(_alone2d if getattr(s, "dim", 3) == 2 else _alone3d)(s, Q, n) for (s, n, _), Q in zip(mix, states)]
    out = {}
    with Batch(hbm_budget=4 * GIB, max_cases=16, use_graph=use_graph, autotune=False) as b:
        ids = {}
        step = 0
        while len(out) < len(mix):
            for k, (s, n, at) in enumerate(mix):
                if at == step and k not in ids.values():
                    ids[b.admit(s, states[k], n)] = k
            b.step()
            step += 1
            for cid in b.done():
                out[ids[cid]] = b.retire(cid)
        st = b.stats()
    for k in range(len(mix)):
        assert np.array_equal(out[k], ref[k]), f"case {k} differs between alone and mix"
    assert st["cell_stages"] == sum(((s.nx - 2) * (s.ny - 2) if getattr(s, "dim", 3) == 2 else s.interior_cells)
                                    * s.stages * n for s, n, _ in mix)


def test_reference_oblique_shock_run():
    """The reference's OS test configuration (SLAU ORDER 6, HR-SLAU2, hybrid
    MUSCL, 257 x 129 points, dt 3e-9 s, 0.1 ms = 33333 steps) in FP64: the
    pre-shock and post-reflected-shock means within the reference test's 3 %
    of its analytic states, and within 0.5 % of the exact regular reflection."""
    spec = c2.os_spec()
    n = int(c2.END_T_REF / c2.DT_REF)
    with Solver2D(spec) as s:
        s.set_state(c2.os_state(spec))
        s.step(n)
        Q = s.get_state()
    pre, post = os_regions(Q)
    u0, p0, T0, rho0 = c2.free_stream()
    for got, want in zip(pre, (rho0, u0, p0)):
        assert abs(got - want) / want < 0.03
    rho3, ux3, p3 = reference_reflected_shock(c2.M0, c2.GAMMA, c2.R_GAS, c2.P_TOT, c2.T_TOT, c2.BETA_REF,
                                              math.pi * 44.1 / 180.0)
    assert abs(post[0] - rho3) / rho3 < 0.03
    assert abs(post[1] - ux3) / ux3 < 0.03
    assert abs(post[2]) / ux3 < 0.03
    assert abs(post[3] - p3) / p3 < 0.03
    ex = c2.reflected_state(c2.BETA_REF)
    for got, want in ((post[0], ex[0]), (post[1], ex[1]), (post[3], ex[3])):
        assert abs(got - want) / want < 0.005
    assert abs(post[2]) / ex[1] < 0.005


@pytest.mark.parametrize("theta_deg", [4.0, 12.0])
def test_fp32_roe_reflection_across_the_deflection_range(theta_deg):
    """The batched ensemble's scheme (Roe + MUSCL, ORDER 6, FP32 bulk) at the
    ends of its deflection range on a 258 x 130 grid over 0.1 ms: the right
    10 % of the domain below half the reflected shock's height there within
    1 % of the exact regular reflection."""
    beta = c2.weak_beta(2.0, math.radians(theta_deg))
    spec = c2.os_spec(258, 130, scheme="Roe", precision="fp32_bulk", beta=beta, cfl=c2.CFL_REF)
    n = int(c2.END_T_REF / spec.dt)
    with Solver2D(spec) as s:
        s.set_state(c2.os_state(spec))
        s.step(n)
        Q = s.get_state()
    rho, u, v, p = primitives(Q)
    y = np.arange(spec.ny) * spec.Ly / (spec.ny - 1)
    rows = y < 0.5 * c2.reflected_shock_height(spec, 0.9 * spec.Lx)
    assert rows.sum() >= 10
    cols = slice(9 * spec.nx // 10, None)
    post = [a[rows][:, cols].mean() for a in (rho, u, v, p)]
    ex = c2.reflected_state(beta)
    for got, want in ((post[0], ex[0]), (post[1], ex[1]), (post[3], ex[3])):
        assert abs(got - want) / want < 0.01, (got, want)
    assert abs(post[2]) / ex[1] < 0.01
