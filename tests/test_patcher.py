"""Patcher and Case routing: the cases of the reference's
ouxsbli/tests/test_patcher.py:38-149, run against this package's own case
files (ouxsbli_amd/case_templates), plus fypp-side rewrites."""
import pytest

from ouxsbli_amd.patcher import patch, _is_param_line, _to_fortran_double
from ouxsbli_amd.case import Case, TEMPLATES, _ALIAS, _VALUE_NORMALIZE
from ouxsbli_amd.fortran_params import read_parameters, read_fypp_macros
from ouxsbli_amd.ic import TaylorGreenIC, UniformIC, RiemannIC, PeriodicBC, ic_to_params, bc_to_params

NSTGV = TEMPLATES / "3D_solver" / "NSTGV"
KHI = TEMPLATES / "3D_solver" / "KHI"


def _line(text, name):
    for l in text.splitlines():
        if _is_param_line(l, name):
            return l
    raise AssertionError(name)


class TestToFortranDouble:
    def test_integer_valued(self):
        assert _to_fortran_double(1600.0) == "1600.0d0"

    def test_fractional(self):
        r = _to_fortran_double(0.03)
        assert r.endswith("d0") and float(r.replace("d0", "e0")) == pytest.approx(0.03)

    def test_exponent(self):
        assert _to_fortran_double(1.5e-7) == "1.5d-07"


class TestScalarPatch3D:
    @pytest.fixture
    def src(self):
        return (NSTGV / "mod_globals.f90").read_text()

    def test_Re(self, src):
        assert "800" in _line(patch(src, {"Re": 800.0}), "Re")

    def test_nx(self, src):
        line = _line(patch(src, {"nx": 65}), "nx")
        assert "65" in line and "513" not in line

    def test_nt_replaces_expression(self, src):
        line = _line(patch(src, {"nt": 500}), "nt")
        assert "500" in line and "int(" not in line

    def test_CFL(self, src):
        assert "0.1" in _line(patch(src, {"CFL": 0.1}), "CFL")

    def test_multiple(self, src):
        out = patch(src, {"Re": 400.0, "nx": 33, "ny": 33, "nz": 33})
        for n in ("nx", "ny", "nz"):
            assert "33" in _line(out, n)

    def test_original_unchanged(self, src):
        patch(src, {"Re": 800.0})
        assert "1600" in src

    def test_comment_kept(self, src):
        out = patch(src, {"threshold": 0.5})
        line = _line(out, "threshold")
        assert "0.5d0" in line and "! hybrid-TVD" in line

    def test_no_append_to_fortran(self, src):
        assert "#:set" not in patch(src, {"NOT_THERE": 3})


class TestFypp:
    @pytest.fixture
    def src(self):
        return (NSTGV / "config.fypp").read_text()

    def test_string_quoted(self, src):
        assert "#:set SCHEME = 'KEEP'" in patch(src, {"SCHEME": "KEEP"})

    def test_bool_and_int(self, src):
        out = patch(src, {"BC_X": True, "ORDER": 4})
        assert "#:set BC_X = True" in out and "#:set ORDER = 4" in out

    def test_append_missing_macro(self, src):
        out = patch(src, {"VISC_ORDER": 4})
        assert out.splitlines()[-1].replace(" ", "") == "#:setVISC_ORDER=4"
        assert read_fypp_macros(out)["VISC_ORDER"] == 4

    def test_derived_macro(self, src):
        assert read_fypp_macros(patch(src, {"ORDER": 4}))["ORDER_IO"] == 1


class TestParameters:
    def test_nstgv_values(self):
        g = read_parameters((NSTGV / "mod_globals.f90").read_text())
        assert g["nx"] == 513 and g["re"] == 1600.0
        assert g["v0"] == pytest.approx(1.25 * (1.4 * 287.03 * 530 * 5 / 9) ** 0.5)
        assert g["nt"] == int(20.0 / (100 * g["dtn"]))

    def test_khi_values(self):
        g = read_parameters((KHI / "mod_globals.f90").read_text())
        assert g["rho1"] == pytest.approx(2.72e3 / (287.15 * 162.7908))


class TestAlias:
    def test_alias(self):
        assert _ALIAS["scheme"] == "SCHEME" and _ALIAS["visc"] == "VISC" and _ALIAS["accuracy"] == "ORDER"

    def test_value_normalize(self):
        assert _VALUE_NORMALIZE["VISC"]["euler"] == "Euler" and _VALUE_NORMALIZE["VISC"]["ns"] == "NS"
        assert _VALUE_NORMALIZE["TVD"]["minmod"] == "tvd" and _VALUE_NORMALIZE["TVD"]["muscl4"] == "hybrid"


class TestCaseBuild:
    def test_routing_and_spec(self, tmp_path):
        c = Case("3D_solver/NSTGV", str(tmp_path / "run"), scheme="keep", visc="euler", accuracy=4, rk="rk4",
                 nx=20, ny=22, nz=24, Re=800.0, np=2, nt=3)
        c.build()
        s = c.spec
        assert (s.scheme, s.visc, s.order, s.rk, s.nx, s.ny, s.nz) == ("KEEP", "Euler", 4, 4, 20, 22, 24)
        assert "#:set SCHEME = 'KEEP'" in (tmp_path / "run" / "config.fypp").read_text()
        assert "800" in _line((tmp_path / "run" / "mod_globals.f90").read_text(), "Re")
        assert (tmp_path / "run" / "build" / "case.json").exists()

    def test_run_before_build(self, tmp_path):
        with pytest.raises(RuntimeError):
            Case("3D_solver/NSTGV", str(tmp_path / "r")).run()

    def test_missing_source(self, tmp_path):
        (tmp_path / "empty").mkdir()
        with pytest.raises(FileNotFoundError):
            Case(str(tmp_path / "empty"), str(tmp_path / "r")).build()

    def test_nonperiodic_rejected(self, tmp_path):
        with pytest.raises(RuntimeError):
            Case("3D_solver/NSTGV", str(tmp_path / "r"), bc_x=True).build()

    def test_ic_specs(self):
        assert ic_to_params(TaylorGreenIC(Ma=0.2))["ic_type"] == 2
        assert ic_to_params({"type": "riemann", "p_l": 2.0})["p_l"] == 2.0
        assert ic_to_params(UniformIC())["ic_type"] == 1
        assert bc_to_params("periodic") == PeriodicBC().to_params()
        with pytest.raises(ValueError):
            ic_to_params({"type": "nope"})


def test_parameters_follow_fortran_integer_semantics():
    """Integer division truncates toward zero and each value takes its
    declared type, as a Fortran compiler evaluates the constant expressions
    of mod_globals.f90; a parameter Case reads that cannot be evaluated
    raises instead of being skipped."""
    text = """
      integer, parameter :: nx = 129
      integer, parameter :: nh = nx/2, unused_tail = 1
      real(8), parameter :: third_int = 1/3
      real(8), parameter :: third = 1.d0/3
      integer, parameter :: neg = -7/2
      real(8), parameter :: inv2 = 2**(-1)
      real(8), parameter :: pi = acos(-1.d0)
      real(8), parameter :: dx = 2.0d0*pi/(nx-1)
      real(8), parameter :: small = 1.e-3/4
      integer, parameter :: trunc = 2.7d0
    """
    g = read_parameters(text, required=())
    assert g["nx"] == 129 and isinstance(g["nx"], int)
    assert g["third_int"] == 0.0 and isinstance(g["third_int"], float)
    assert g["third"] == 1.0 / 3.0
    assert g["neg"] == -3          # toward zero, not floor
    assert g["inv2"] == 0.0        # 2**(-1) is integer 0 in Fortran
    assert g["dx"] == 2.0 * g["pi"] / 128
    assert g["small"] == 1.0e-3 / 4
    assert g["trunc"] == 2
    with pytest.raises(ValueError):
        read_parameters("integer, parameter :: nx = undefined_function(3)\n")


def test_reference_import_names():
    """The reference's scripts import from ``ouxsbli`` (tutorials/test.py:3-5,
    ouxsbli/tests/test_st.py:10); those names resolve to the ouxsbli_amd package."""
    import ouxsbli
    from ouxsbli import Case as C2
    from ouxsbli.patcher import patch as p2
    from ouxsbli.ic import TaylorGreenIC as T2
    assert C2 is Case and p2 is patch and T2 is TaylorGreenIC
    assert set(ouxsbli.__all__) == {"Case", "patch", "UniformIC", "TaylorGreenIC", "RiemannIC", "PeriodicBC"}
