"""Non-periodic directions on the GPU: slip walls and zero-gradient ends.

The FP64 kernels with the boundary fills (ghost primitives after every stage,
the Ducros sensor of the inner ghost layer, the state's ghost layers at the
end of a call) equal the oracle bit for bit, whose fill restates
3D_solver/STZ/set.f90:77-138 (tests/test_bc.py); Sod's problem between slip
walls (the STZ case on one domain) meets the exact solution in both forms.
"""
import numpy as np
import pytest

from ouxsbli_amd import cases
from ouxsbli_amd.batch import Batch
from ouxsbli_amd.cases import CaseSpec
from ouxsbli_amd.solver import Solver
from oracle_lib import Oracle
from test_bc import STZ_L1, _field, stz_l1

pytestmark = pytest.mark.gpu

CASES = {
    # name: (interior (nx, ny, nz), order, scheme, tvd, variant, rk, bc, steps)
    "hybrid6-x-slip-z-open": ((130, 20, 72), 6, "Hybrid", "hybrid", "HRSLAU2", 3,
                              ("slip", "periodic", "extrapolate"), 3),
    "slau6-muscl-y-z-slip": ((60, 72, 40), 6, "SLAU", "tvd", "HRSLAU2", 3, ("extrapolate", "slip", "slip"), 3),
    "keep4-rk4-walls": ((20, 24, 28), 4, "KEEP", "none", "HRSLAU2", 4, ("slip", "slip", "slip"), 4),
    "slau4-y-open": ((36, 30, 20), 4, "SLAU", "tvd", "SLAU", 3, ("periodic", "extrapolate", "periodic"), 4),
    # Navier-Stokes: the supersonic TGV's scheme between slip walls in y and z, and
    # ME4Base / central-4 variants
    "slau6-ns6-y-z-slip": ((130, 40, 72), 6, "SLAU", "hybrid", "HRSLAU2", 3, ("periodic", "slip", "slip"), 2,
                           "NS", 6, "Central"),
    "hybrid6-me4-walls": ((60, 30, 24), 6, "Hybrid", "hybrid", "HRSLAU2", 3, ("slip", "extrapolate", "slip"), 2,
                          "NS", 6, "ME4Base"),
    "keep4-ns4-x-open": ((40, 20, 24), 4, "KEEP", "none", "HRSLAU2", 4, ("extrapolate", "periodic", "slip"), 3,
                         "NS", 4, "Central"),
}


def _case(name, precision="fp64"):
    (nx, ny, nz), order, scheme, tvd, variant, rk, bc, steps, *visc = CASES[name]
    visc, vorder, stencil = visc if visc else ("Euler", 0, "Central")
    spec = CaseSpec(nx=nx + order, ny=ny + order, nz=nz + order, order=order, scheme=scheme, visc=visc,
                    visc_order=vorder, visc_stencil=stencil, tvd=tvd, slau_variant=variant, rk=rk,
                    precision=precision, threshold=0.4, dt=1e-3, Lx=1.0, Ly=0.9, Lz=1.1, bc=bc)
    h = order // 2
    Q = np.zeros(spec.state_shape())
    Q[h:-h, h:-h, :, h:-h] = _field((nz, ny, nx), seed=len(name))
    o = Oracle(spec)
    o.set_qj(Q)  # the fill copies and negates: the same on Q as on QJ
    o.set_bc()
    Q = o.get_qj()
    return spec, Q, steps


@pytest.mark.parametrize("name", list(CASES))
def test_walls_bitwise_vs_oracle(name):
    spec, Q, steps = _case(name)
    print(f"boundary case {name} {spec.nx}x{spec.ny}x{spec.nz} bc {spec.bc}", flush=True)
    o = Oracle(spec)
    o.set_state(Q)
    Ro = o.rhs()
    o.step(steps)
    with Solver(spec) as s:
        s.set_state(Q)
        Rg = s.rhs()
        s.step(steps)
        QJg = s.get_qj()
        Rg2 = s.rhs()
    h = spec.order // 2
    sl = (slice(h, -h), slice(h, -h), slice(None), slice(h, -h))
    np.testing.assert_array_equal(Rg[sl], Ro[sl])
    np.testing.assert_array_equal(QJg, o.get_qj())  # ghost layers included
    np.testing.assert_array_equal(Rg2[sl], o.rhs()[sl])


@pytest.mark.parametrize("name", ["hybrid6-x-slip-z-open", "slau6-ns6-y-z-slip"])
def test_walls_fp32_bulk_close_to_oracle(name):
    spec, Q, steps = _case(name, "fp32_bulk")
    o = Oracle(CaseSpec(**{**spec.__dict__, "precision": "fp64"}))
    o.set_state(Q)
    o.step(steps)
    with Solver(spec) as s:
        s.set_state(Q)
        s.step(steps)
        Qg = s.get_state()
        ke_g = s.diagnostics()["ke"]
    Qo = o.get_state()
    for l in range(5):
        scale = np.abs(Qo[:, :, l]).max()
        assert np.abs(Qg[:, :, l] - Qo[:, :, l]).max() <= 2e-4 * scale, l
    assert abs(ke_g - o.diagnostics()["ke"]) <= 1e-5 * abs(o.diagnostics()["ke"])


@pytest.mark.parametrize("precision", ["fp64", "fp32_bulk"])
def test_stz_sod_between_walls(precision):
    spec = cases.stz_spec(256, 4, precision=precision)
    n = int(0.1 / spec.dt)
    with Solver(spec) as s:
        s.set_state(cases.stz_state(spec))
        s.step(n)
        Q = s.get_state()
    l1_rho, l1_p = stz_l1(Q, spec, n * spec.dt)
    assert l1_rho < STZ_L1 and l1_p < STZ_L1, (precision, l1_rho, l1_p)


def test_stz_fp64_equals_oracle():
    spec = cases.stz_spec(64, 4)
    Q = cases.stz_state(spec)
    o = Oracle(spec)
    o.set_state(Q)
    o.step(40)
    with Solver(spec) as s:
        s.set_state(Q)
        s.step(40)
        np.testing.assert_array_equal(s.get_qj(), o.get_qj())


@pytest.mark.parametrize("kw", [dict(scheme="Roe"), dict(visc="LES", visc_order=4)])
def test_walls_refused_where_not_built(kw):
    spec = cases.stz_spec(nz=32, nt=16, order=4)
    for k, v in kw.items():
        setattr(spec, k, v)
    with pytest.raises(RuntimeError):
        Solver(spec)


@pytest.mark.parametrize("precision", ["fp64", "fp32_bulk"])
@pytest.mark.parametrize("use_graph", [True, False])
def test_walls_batched_equal_alone(precision, use_graph):
    """Cases with non-periodic directions in the batch engine (their
    boundary fills follow their group's stage, captured in the step graphs)
    beside periodic cases of the same specialisation: each ends bit-identical
    to its ox_step run."""
    mix = []
    for name in ("slau6-ns6-y-z-slip", "keep4-rk4-walls", "hybrid6-me4-walls"):
        spec, Q, _ = _case(name, precision)
        mix.append((spec, Q, 3))
    stz = cases.stz_spec(64, 4, precision=precision)
    mix.append((stz, cases.stz_state(stz), 5))
    per = cases.tgv_spec(20, 6, Ma=1.25, cfl=0.3, cfl_mode="acoustic", scheme="SLAU", visc="NS", visc_order=6,
                         rk=3, tvd="hybrid", precision=precision)
    mix.append((per, cases.initial_state(per), 4))
    ref = []
    for spec, Q, n in mix:
        with Solver(spec) as s:
            s.set_state(Q)
            s.step(n)
            ref.append(s.get_state())
    out = {}
    with Batch(hbm_budget=8 << 30, max_cases=8, use_graph=use_graph, autotune=False) as b:
        ids = {b.admit(spec, Q, n): k for k, (spec, Q, n) in enumerate(mix)}
        while len(out) < len(mix):
            b.run(10)
            for cid in b.done():
                out[ids[cid]] = b.retire(cid)
    for k in range(len(mix)):
        np.testing.assert_array_equal(out[k], ref[k], err_msg=f"case {k}")
