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
