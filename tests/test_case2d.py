"""Parameter routing of the drop-in Case for the 2-D oblique-shock case
(no GPU): the reference's OS test parameters reach the 2-D specification."""
import math

from ouxsbli_amd import Case
from ouxsbli_amd.cases2d import CaseSpec2D


def test_os_parameters_route_to_the_2d_spec(tmp_path):
    c = Case("2D_solver/OS", str(tmp_path / "os"), scheme="roe", accuracy=4, visc="euler", tvd="muscl4",
             nx=131, ny=67, Lx=4e-3, Ly=1.5e-3, gamma=1.4, R=287.0, M0=2.5, p_tot=90e3, T_tot=300.0,
             beta=math.radians(30.0), dt=5e-9, np=3, nt=7, slau="slau")
    c.build()
    s = c.spec
    assert isinstance(s, CaseSpec2D)
    assert (s.nx, s.ny, s.order, s.scheme, s.tvd, s.slau_variant) == (131, 67, 4, "Roe", "hybrid", "SLAU")
    assert (s.Lx, s.Ly, s.dt, s.M, s.p_tot, s.T_tot, s.R) == (4e-3, 1.5e-3, 5e-9, 2.5, 90e3, 300.0, 287.0)
    assert abs(s.beta - math.radians(30.0)) < 1e-11  # the patcher writes 10 significant digits
    assert s.top_split == 13
    assert (c.globals["np"], c.globals["nt"]) == (3, 7)
    cfg = s.to_ox()
    assert cfg.nx == 131 and cfg.top_split == 13 and cfg.bc == 1


def test_os_defaults_are_the_reference_case(tmp_path):
    c = Case("2D_solver/OS", str(tmp_path / "os"))
    c.build()
    s = c.spec
    assert (s.nx, s.ny, s.order, s.scheme, s.tvd, s.slau_variant, s.dt) == (257, 129, 6, "SLAU", "hybrid",
                                                                             "HRSLAU2", 3e-9)
    assert abs(s.beta - math.pi * 37.2 / 180.0) < 1e-15
