"""``Case`` -- one simulation run, the drop-in for the reference's
``ouxsbli.Case`` (ouxsbli/case.py:68-256).

Same constructor, aliases, value normalisation and parameter routing:
``build()`` copies the case directory into ``workdir`` and patches
``config.fypp`` / ``mod_globals.f90``; then, where the reference runs
cmake/make to compile one executable, it reads the patched files into an
``ox_config`` and checks that the gfx950 library holds that specialisation.
``run()`` drives the library in-process through ctypes instead of
``mpirun a.out`` and writes the same run directory: ``data/Q%05d.vtr``,
``data/kinetic_energy.d``, ``data/entropy.d`` and ``recal/Q00001.dat``.
Two-dimensional cases (``dimension = 2`` in mod_globals.f90, the
2D_solver/OS oblique-shock case) run on the 2-D path and write what the
reference's 2-D executable writes: ``data/Q%05d.vtr`` (one z plane) and
``recal/Q00001.dat``, ``recal/x.dat``, ``recal/y.dat``.
"""
from __future__ import annotations

import json
import pathlib
import shutil
from typing import Any

import numpy as np

from . import io_vtr
from .cases import CaseSpec, cell_centres
from .cases2d import CaseSpec2D
from .fortran_params import read_fypp_macros, read_parameters
from .ic import CASE_BOUNDARIES, bc_to_params, device_ic, ic_to_params, initial_state
from .patcher import _is_macro_line, patch

_ALIAS = {
    "scheme": "SCHEME", "visc": "VISC", "accuracy": "ORDER", "visc_order": "VISC_ORDER", "tvd": "TVD",
    "slau": "SLAU_VARIANT", "rescale": "RESCALE", "recal": "RESTART", "rk": "RK", "gpumpi": "GPUMPI",
    "bc_x": "BC_X", "bc_y": "BC_Y", "commz": "COMMZ",
    # build-specific (no counterpart in the reference): arithmetic of the RHS
    "precision": "PRECISION", "visc_stencil": "VISC_STENCIL",
}
_VALUE_NORMALIZE = {
    "SCHEME": {"keep": "KEEP", "slau": "SLAU", "roe": "Roe", "hybrid": "Hybrid"},
    "VISC": {"euler": "Euler", "ns": "NS", "les": "LES"},
    "TVD": {"none": "none", "tvd": "tvd", "hybrid": "hybrid", "minmod": "tvd", "muscl4": "hybrid"},
    "SLAU_VARIANT": {"slau": "SLAU", "hrslau2": "HRSLAU2"},
}
_RK_NORMALIZE = {"tvd_rk3": 3, "rk3": 3, "rk4": 4, "classical_rk4": 4}

TEMPLATES = pathlib.Path(__file__).resolve().parent / "case_templates"


class Case:
    """Manage one run: copy source -> patch -> specialise -> run on the GPU."""

    def __init__(self, source: str, workdir: str, **params: Any) -> None:
        src = pathlib.Path(source)
        self.source = src.resolve() if src.is_absolute() or src.exists() else (TEMPLATES / source).resolve()
        self.workdir = pathlib.Path(workdir).resolve()
        self.build_dir = self.workdir / "build"
        raw = dict(params)
        ic_spec = raw.pop("ic", None)
        bc_spec = raw.pop("bc", None)
        if ic_spec is not None:
            raw.update(ic_to_params(ic_spec))
        if bc_spec is not None:
            raw.update(bc_to_params(bc_spec))
        self.params = {_ALIAS.get(k.lower(), k): v for k, v in raw.items()}
        self._built = False
        self.spec: CaseSpec | CaseSpec2D | None = None
        self.globals: dict = {}
        self.macros: dict = {}

    # -- setup: copy + patch (case.py:111-152 of the reference) --------------
    def _setup(self) -> None:
        if self.workdir.exists():
            shutil.rmtree(self.workdir)
        shutil.copytree(self.source, self.workdir,
                        ignore=shutil.ignore_patterns("build", "CMakeCache.txt", "CMakeFiles", "*.cmake"))
        cfg_path = self.workdir / "config.fypp"
        glob_path = self.workdir / "mod_globals.f90"
        if not cfg_path.exists():
            raise FileNotFoundError(f"config.fypp not found in {self.source}")
        if not glob_path.exists():
            raise FileNotFoundError(f"mod_globals.f90 not found in {self.source}")
        cfg_text, glob_text = cfg_path.read_text(), glob_path.read_text()
        cfg_lines = cfg_text.splitlines()
        to_cfg, to_glob = {}, {}
        for k, v in self.params.items():
            if k == "RK" and isinstance(v, str):
                v = _RK_NORMALIZE.get(v.lower(), v)
            if any(_is_macro_line(line, k) for line in cfg_lines):
                if isinstance(v, str) and k in _VALUE_NORMALIZE:
                    v = _VALUE_NORMALIZE[k].get(v.lower(), v)
                to_cfg[k] = v
            else:
                to_glob[k] = v
        cfg_path.write_text(patch(cfg_text, to_cfg))
        glob_path.write_text(patch(glob_text, to_glob))

    # -- build ---------------------------------------------------------------
    def build(self) -> None:
        """Apply patches, read the case into an ox_config and select the
        prebuilt gfx950 specialisation (raises if none matches)."""
        self._setup()
        self.macros = read_fypp_macros((self.workdir / "config.fypp").read_text())
        self.globals = read_parameters((self.workdir / "mod_globals.f90").read_text())
        # parameters that were routed to mod_globals but are not declared there
        # (e.g. IC spec fields) still reach the initial-condition generator
        for k, v in self.params.items():
            if k not in self.macros:
                self.globals.setdefault(k.lower(), v)
        self.spec = self._make_spec()
        self.build_dir.mkdir(exist_ok=True)
        (self.build_dir / "case.json").write_text(json.dumps(
            {"source": str(self.source), "macros": {k: v for k, v in self.macros.items()},
             "spec": {k: getattr(self.spec, k) for k in self.spec.__dataclass_fields__}}, indent=1, default=str))
        self._built = True

    def _make_spec(self):
        m, g = self.macros, self.globals
        if int(g.get("dimension", 3)) == 2:
            return self._make_spec2d()
        # BC_*=True selects the reference's order-fallback flux kernels, whose
        # faces next to interior cells are the periodic kernels' own; what
        # differs is the case's ghost fill (its set_bc), built for the cases
        # of CASE_BOUNDARIES.  Their z slabs (COMMZ) are one domain here.
        bc = CASE_BOUNDARIES.get(self.source.name)
        if bc is None:
            for flag in ("BC_X", "BC_Y", "BC_Z"):
                if m.get(flag, False):
                    raise RuntimeError(f"{flag}=True: the boundary fill of case {self.source.name!r} is not built "
                                       "(non-periodic fills: " + ", ".join(CASE_BOUNDARIES) + ")")
            if m.get("COMMZ", False):
                raise RuntimeError("COMMZ=True: one GPU holds the whole grid; z-slab decomposition is not built")
            bc = ("periodic",) * 3
        if m.get("RESCALE", False):
            raise RuntimeError("RESCALE=True (SBLI inflow recycling) is not part of this library")
        order = int(m.get("ORDER", 6))
        spec = CaseSpec(
            nx=int(g["nx"]), ny=int(g["ny"]), nz=int(g["nz"]), order=order, scheme=str(m.get("SCHEME", "KEEP")),
            visc=str(m.get("VISC", "Euler")), visc_order=int(m.get("VISC_ORDER", order)),
            visc_stencil=str(m.get("VISC_STENCIL", "Central")), tvd=str(m.get("TVD", "none")),
            slau_variant=str(m.get("SLAU_VARIANT", "HRSLAU2")), rk=int(m.get("RK", 3)),
            precision=str(m.get("PRECISION", "fp64")), gamma=float(g.get("gamma", 1.4)), R=float(g.get("r", 287.03)),
            Pr=float(g.get("pr", 0.71)), Prt=float(g.get("prt", 0.9)), threshold=float(g.get("threshold", 0.4)),
            dt=float(g["dt"]), Lx=float(g["lx"]), Ly=float(g["ly"]), Lz=float(g["lz"]), bc=bc)
        if spec.visc == "Euler":
            spec.visc_order = 0
        return spec

    def _make_spec2d(self) -> CaseSpec2D:
        """The 2-D oblique-shock case: non-periodic x and y (BC_X, BC_Y) with
        the boundary conditions of 2D_solver/OS/set.f90:48-86."""
        m, g = self.macros, self.globals
        if self.source.name != "OS":
            raise RuntimeError(f"2-D case {self.source.name!r}: only the oblique-shock case (OS) is built")
        if not (m.get("BC_X", True) and m.get("BC_Y", True)):
            raise RuntimeError("the 2-D path runs the oblique-shock boundaries (BC_X = BC_Y = True)")
        if str(m.get("VISC", "Euler")) != "Euler":
            raise RuntimeError("the 2-D path is inviscid (VISC = 'Euler')")
        if int(m.get("RK", 3)) != 3:
            raise RuntimeError("the 2-D path advances with RK3")
        return CaseSpec2D(
            nx=int(g["nx"]), ny=int(g["ny"]), order=int(m.get("ORDER", 6)), scheme=str(m.get("SCHEME", "SLAU")),
            tvd=str(m.get("TVD", "hybrid")), slau_variant=str(m.get("SLAU_VARIANT", "HRSLAU2")), rk=3,
            precision=str(m.get("PRECISION", "fp64")), gamma=float(g.get("gamma", 1.4)),
            R=float(g.get("r", 287.03)), threshold=float(g.get("threshold", 0.4)), dt=float(g["dt"]),
            Lx=float(g["lx"]), Ly=float(g["ly"]), M=float(g.get("m0", 2.0)), p_tot=float(g.get("p_tot", 100e3)),
            T_tot=float(g.get("t_tot", 295.0)), beta=float(g["beta"]))

    def _run2d(self) -> None:
        """RungeKutta of the 2-D executable (2D_solver/src/calc_time_dev.f90.fypp:74-102):
        np outputs of nt steps, one .vtr per output (print_vtk_2D,
        src/print.f90:351-361), then the restart and grid files
        (2D_solver/src/main.f90:80-99)."""
        from .solver2d import Solver2D
        spec, g = self.spec, self.globals
        data = self.workdir / "data"
        data.mkdir(exist_ok=True)
        recal = self.workdir / "recal"
        recal.mkdir(exist_ok=True)
        nt, np_ = int(g.get("nt", 1)), int(g.get("np", 1))
        offset = int(g.get("step_offset", 0))
        # a fresh run builds OS set_init on the device; a restart uploads its record
        Q = None
        if bool(self.macros.get("RESTART", False)):
            Q = io_vtr.read_restart(recal / "Q00001.dat", spec.state_shape())
        # set_grid (OS/set.f90:8-26): running sums of the uniform spacing
        x = np.zeros(spec.nx)
        y = np.zeros(spec.ny)
        dx0, dy0 = spec.Lx / float(spec.nx - 1), spec.Ly / float(spec.ny - 1)
        for i in range(1, spec.nx):
            x[i] = x[i - 1] + dx0
        for j in range(1, spec.ny):
            y[j] = y[j - 1] + dy0
        z = np.zeros(1)
        with Solver2D(spec) as s:
            if Q is None:
                s.init_state()
            else:
                s.set_state(Q)

            def output(step):
                rho, p, vel = io_vtr.primitives_for_print2d(s.get_state(), spec.gamma)
                io_vtr.write_vtr(data / f"Q{step + offset:05d}.vtr", x, y, z, rho, p, vel)

            output(0)
            for t2 in range(1, np_ + 1):
                s.step(nt)
                output(t2)
            Qf = s.get_state()
        if not np.isfinite(Qf).all():
            raise RuntimeError("non-finite state at the end of the 2-D run")
        io_vtr.write_restart(recal / "Q00001.dat", Qf)
        (recal / "x.dat").write_bytes(np.ascontiguousarray(x, dtype="<f8").tobytes())
        (recal / "y.dat").write_bytes(np.ascontiguousarray(y, dtype="<f8").tobytes())

    # -- run -------------------------------------------------------------------
    def run(self, nranks: int = 2) -> None:
        """Advance np outputs x nt steps on one GPU, writing the run directory.
        ``nranks`` is accepted for compatibility (the reference's I/O rank is
        replaced by in-process writes)."""
        if not self._built:
            raise RuntimeError("Must call build() before run()")
        if isinstance(self.spec, CaseSpec2D):
            self._run2d()
            return
        from .solver import Solver
        spec = self.spec
        g = self.globals
        data = self.workdir / "data"
        data.mkdir(exist_ok=True)
        (self.workdir / "recal").mkdir(exist_ok=True)
        nt, np_ = int(g.get("nt", 1)), int(g.get("np", 1))
        restart = bool(self.macros.get("RESTART", False))
        # the case's set_init: on the device where the library has it (no host
        # array, no upload), else on the host; a restart uploads its record
        Q, dev_ic = None, None
        if restart:
            Q = io_vtr.read_restart(self.workdir / "recal" / "Q00001.dat", spec.state_shape())
        else:
            dev_ic = device_ic(self.source.name, spec, g)
            if dev_ic is None:
                Q = initial_state(self.source.name, spec, g)
        x = cell_centres(spec.nx, spec.Lx, spec.order)
        y = cell_centres(spec.ny, spec.Ly, spec.order)
        z = cell_centres(spec.nz, spec.Lz, spec.order)
        diag = io_vtr.Diagnostics(data)
        offset = int(g.get("step_offset", 0))
        with Solver(spec) as s:
            if dev_ic is not None:
                s.init_state(dev_ic)
            else:
                s.set_state(Q)
            jac = self.jacobian()

            def output(step):
                QJ = s.get_qj()
                rho, p, vel = io_vtr.primitives_for_print(QJ, jac, spec.gamma)
                ke, ent = io_vtr.ke_entropy(rho, p, vel, spec.order, spec.gamma)
                diag.record(step, nt * step * spec.dt, ke, ent)
                io_vtr.write_vtr(data / f"Q{step + offset:05d}.vtr", x, y, z, rho, p, vel)
                with open(data / "diagnostics_fp64.d", "a") as f:
                    d = s.diagnostics()
                    f.write(f"{nt * step * spec.dt:.16e} {d['ke']:.16e} {d['entropy']:.16e} {d['enstrophy']:.16e} "
                            f"{d['max_wave_speed']:.16e}\n")

            output(0)
            for t2 in range(1, np_ + 1):
                s.step(nt)
                ok, bad, cnt = s.check_state()
                if not ok:
                    raise RuntimeError(f"non-physical state after {t2 * nt} steps: {cnt} cells, first at {bad}")
                output(t2)
            io_vtr.write_restart(self.workdir / "recal" / "Q00001.dat", s.get_state())

    def jacobian(self) -> float:
        s = self.spec
        dx = s.Lx / (s.nx - s.order)
        dy = s.Ly / (s.ny - s.order)
        dz = s.Lz / (s.nz - s.order)
        return 8.0 / ((dx + dx) * (dy + dy) * 2.0 * dz)

    def __repr__(self) -> str:
        return f"Case(source={self.source.name!r}, workdir={str(self.workdir)!r}, params={self.params})"
