"""End to end through the drop-in Case API on the GPU: the run directory the
reference writes, and restart round trips."""
import numpy as np
import pytest

from ouxsbli_amd import Case
from ouxsbli_amd import io_vtr
from ouxsbli_amd.cases import cell_centres

pytestmark = pytest.mark.gpu


def _case(tmp, name="run", **kw):
    p = dict(nx=22, ny=22, nz=22, np=2, nt=2, accuracy=6)
    p.update(kw)
    c = Case("3D_solver/NSTGV", str(tmp / name), **p)
    c.build()
    return c


def test_run_writes_reference_files(tmp_path):
    c = _case(tmp_path)
    c.run()
    d = tmp_path / "run" / "data"
    for s in range(3):
        assert (d / f"Q{s:05d}.vtr").exists()
    x, y, z, rho, p, vel = io_vtr.read_vtr(d / "Q00002.vtr")
    assert rho.shape == (22, 22, 22) and vel.shape == (22, 22, 22, 3)
    np.testing.assert_allclose(x, cell_centres(22, c.spec.Lx, 6).astype(np.float32))
    assert np.all(rho > 0) and np.all(p > 0)
    ke = [l.split() for l in (d / "kinetic_energy.d").read_text().splitlines()]
    assert len(ke) == 3 and float(ke[0][2]) == 1.0
    ent = (d / "entropy.d").read_text().splitlines()
    assert len(ent) == 3 and float(ent[0].split()[1]) == 0.0
    assert (tmp_path / "run" / "recal" / "Q00001.dat").stat().st_size == 22 * 5 * 22 * 22 * 8


def test_restart_continues_bitwise(tmp_path):
    a = _case(tmp_path, "a", np=1, nt=4)
    a.run()
    full = io_vtr.read_restart(tmp_path / "a" / "recal" / "Q00001.dat", a.spec.state_shape())
    b = _case(tmp_path, "b", np=1, nt=2)
    b.run()
    half = (tmp_path / "b" / "recal" / "Q00001.dat").read_bytes()
    c = _case(tmp_path, "c", np=1, nt=2, recal=True)
    (tmp_path / "c" / "recal").mkdir(exist_ok=True)
    (tmp_path / "c" / "recal" / "Q00001.dat").write_bytes(half)
    c.run()
    again = io_vtr.read_restart(tmp_path / "c" / "recal" / "Q00001.dat", c.spec.state_shape())
    # the restart file holds J*QJ, re-divided by J on load: equal to ~1 ulp
    np.testing.assert_allclose(again, full, rtol=1e-13, atol=0)


def test_oblique_shock_case_like_the_reference_test(tmp_path):
    """The reference's OS test (ouxsbli/tests/test_os.py) through the Case API:
    same parameters, one output after 0.1 ms, region means of the written
    .vtr within 3 % of the analytic free stream and reflected-shock states."""
    import math
    from ouxsbli_amd import cases2d as c2
    from test_oracle2d import reference_reflected_shock
    nx, ny, endT, dt = 257, 129, 0.1e-3, 3.0e-9
    case = Case("2D_solver/OS", str(tmp_path / "os"), scheme="slau", accuracy=6, visc="euler", tvd="hybrid",
                nx=nx, ny=ny, Lx=5e-3, Ly=2e-3, gamma=1.4, Pr=0.72, R=287.03, M0=2.0, p_tot=100e3, T_tot=295.0,
                beta=math.pi * 37.2 / 180.0, dt=dt, np=1, nt=int(endT / dt))
    case.build()
    case.run(nranks=2)
    d = tmp_path / "os" / "data"
    assert (d / "Q00000.vtr").exists()
    x, y, z, rho, p, vel = io_vtr.read_vtr(d / "Q00001.vtr")
    ni, nj, nk = len(x), len(y), len(z)
    assert (ni, nj, nk) == (nx, ny, 1)
    u, v = vel[..., 0], vel[..., 1]
    u0, p0, T0, rho0 = c2.free_stream()
    pre = [a[0, nj // 2:, : ni // 10].astype(float).mean() for a in (rho, u, p)]
    for got, want in zip(pre, (rho0, u0, p0)):
        assert abs(got - want) / want < 0.03
    rho3, ux3, p3 = reference_reflected_shock(2.0, 1.4, 287.03, 100e3, 295.0, math.pi * 37.2 / 180.0,
                                              math.pi * 44.1 / 180.0)
    post = [a[0, : nj // 3, 9 * ni // 10:].astype(float).mean() for a in (rho, u, v, p)]
    assert abs(post[0] - rho3) / rho3 < 0.03
    assert abs(post[1] - ux3) / ux3 < 0.03
    assert abs(post[2]) / abs(ux3) < 0.03
    assert abs(post[3] - p3) / p3 < 0.03
    recal = tmp_path / "os" / "recal"
    assert (recal / "Q00001.dat").stat().st_size == nx * 4 * ny * 8
    assert (recal / "x.dat").stat().st_size == nx * 8


def test_stz_case_between_walls(tmp_path):
    """The STZ case (Sod's problem along z between slip walls, zero-gradient
    ends) through the Case API: its 50 outputs, and the final state against
    the exact solution, as on the solver (tests/test_gpu_bc.py)."""
    from ouxsbli_amd import cases
    from test_bc import STZ_L1, stz_l1
    c = Case("3D_solver/STZ", str(tmp_path / "stz"))
    c.build()
    assert c.spec.bc == ("slip", "slip", "extrapolate")
    c.run()
    d = tmp_path / "stz" / "data"
    n_out = int(c.globals["np"])
    assert (d / f"Q{n_out:05d}.vtr").exists()
    assert len((d / "kinetic_energy.d").read_text().splitlines()) == n_out + 1
    Q = io_vtr.read_restart(tmp_path / "stz" / "recal" / "Q00001.dat", c.spec.state_shape())
    steps = n_out * int(c.globals["nt"])
    l1_rho, l1_p = stz_l1(Q, c.spec, steps * c.spec.dt)
    assert l1_rho < STZ_L1 and l1_p < STZ_L1, (l1_rho, l1_p)
    assert cases.stz_spec(256, 4).dt == c.spec.dt


def test_etgv_case_like_the_reference_test(tmp_path):
    """The reference's ETGV test (ouxsbli/tests/test_etgv.py:15-60) through
    the Case API with its own parameters (KEEP ORDER 2, Euler, RK4, 66^3,
    20000 steps of dt 0.01, one output) and its thresholds on the .d files it
    reads: |(s0-s)/s0| < 3e-3 and 0.75 < ke/ke0 < 1.01."""
    case = Case("3D_solver/ETGV", str(tmp_path / "etgv"), visc="euler", scheme="keep", accuracy=2, rk="rk4",
                nx=66, ny=66, nz=66, nt=20000, np=1, dt=0.01e0)
    case.build()
    case.run(nranks=2)
    d = tmp_path / "etgv" / "data"
    ent = np.loadtxt(d / "entropy.d")
    ke = np.loadtxt(d / "kinetic_energy.d")
    assert ent.shape == (2, 2) and ke.shape == (2, 3)
    assert abs(ent[-1, -1]) < 3.0e-3, ent
    assert 0.75 < ke[-1, -1] < 1.0 + 0.01, ke


def test_khi_case_from_device_state(tmp_path):
    """Case("3D_solver/KHI") builds the KHI set_init on the device
    (OX_IC_KHI); its run agrees with the same run from the host form of the
    state to 1e-12 of each variable's scale (the device's sin rounds within an
    ulp of the host's)."""
    from ouxsbli_amd import ic
    from ouxsbli_amd.solver import Solver
    c = Case("3D_solver/KHI", str(tmp_path / "khi"), nx=24, ny=24, nz=24, np=1, nt=4)
    c.build()
    assert ic.device_ic("KHI", c.spec, c.globals) is not None
    c.run()
    got = io_vtr.read_restart(tmp_path / "khi" / "recal" / "Q00001.dat", c.spec.state_shape())
    with Solver(c.spec) as s:
        s.set_state(ic.khi_state(c.spec, c.globals))
        s.step(4)
        ref = s.get_state()
    for l in range(5):
        scale = np.abs(ref[:, :, l, :]).max()
        assert np.abs(got[:, :, l, :] - ref[:, :, l, :]).max() <= 1e-12 * scale, l
