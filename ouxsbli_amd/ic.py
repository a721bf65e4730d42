"""Initial- and boundary-condition specs for :class:`ouxsbli_amd.Case`.

Same classes, fields and parameter names as the reference's ``ouxsbli.ic``
(ouxsbli/ic.py:20-198): each spec flattens to ``mod_globals.f90`` parameters
with ``to_params()``; ``ic_to_params`` / ``bc_to_params`` accept instances,
dicts with a ``"type"`` key, or (for BCs) the string ``"periodic"``.  Here the
flattened parameters also select the state generator of :func:`initial_state`.
"""
from __future__ import annotations

from dataclasses import dataclass, fields

import numpy as np

IC_UNIFORM = 1
IC_TAYLOR_GREEN = 2
IC_RIEMANN = 3


@dataclass
class UniformIC:
    rho: float = 1.0
    u: float = 0.0
    v: float = 0.0
    w: float = 0.0
    p: float = 0.71429

    def to_params(self) -> dict:
        return {"ic_type": IC_UNIFORM, "rho0": self.rho, "u0": self.u, "v0": self.v, "w0": self.w, "p0": self.p}


@dataclass
class TaylorGreenIC:
    """u = Ma sin x cos y cos z, v = -Ma cos x sin y cos z, w = 0,
    p = p0 + rho0 Ma^2/16 (cos 2x + cos 2y)(cos 2z + 2) on [0, 2 pi)^3."""
    Ma: float = 0.1
    rho0: float = 1.0
    p0: float = 0.71429

    def to_params(self) -> dict:
        return {"ic_type": IC_TAYLOR_GREEN, "Ma": self.Ma, "rho0": self.rho0, "p0": self.p0}


@dataclass
class RiemannIC:
    """Left/right states split at x = x0 * Lx (homogeneous in y, z)."""
    rho_l: float = 1.0
    u_l: float = 0.0
    p_l: float = 1.0
    rho_r: float = 0.125
    u_r: float = 0.0
    p_r: float = 0.1
    x0: float = 0.5

    def to_params(self) -> dict:
        return {"ic_type": IC_RIEMANN, "rho_l": self.rho_l, "u_l": self.u_l, "p_l": self.p_l, "rho_r": self.rho_r,
                "u_r": self.u_r, "p_r": self.p_r, "x0_frac": self.x0}


@dataclass
class PeriodicBC:
    """Periodic in all three directions (BC_X = BC_Y = BC_Z = False)."""

    def to_params(self) -> dict:
        return {"BC_X": False, "BC_Y": False, "BC_Z": False}


def _from_dict(d: dict):
    d = dict(d)
    kind = str(d.pop("type", "uniform")).lower()
    cls = {"uniform": UniformIC, "tgv": TaylorGreenIC, "taylor_green": TaylorGreenIC, "riemann": RiemannIC}.get(kind)
    if cls is None:
        raise ValueError(f"Unknown IC type: {kind!r}.  Supported: 'uniform', 'tgv', 'riemann'.")
    names = {f.name for f in fields(cls)}
    return cls(**{k: v for k, v in d.items() if k in names})


def ic_to_params(ic) -> dict:
    if isinstance(ic, dict):
        ic = _from_dict(ic)
    if hasattr(ic, "to_params"):
        return ic.to_params()
    raise TypeError(f"Unsupported IC specification: {type(ic)!r}")


def bc_to_params(bc) -> dict:
    if isinstance(bc, str):
        if bc.lower() == "periodic":
            return PeriodicBC().to_params()
        raise ValueError(f"Unknown BC type string: {bc!r}.  Supported: 'periodic'.")
    if isinstance(bc, dict):
        kind = str(bc.get("type", "periodic")).lower()
        if kind == "periodic":
            return PeriodicBC().to_params()
        raise ValueError(f"Unknown BC type: {kind!r}.")
    if hasattr(bc, "to_params"):
        return bc.to_params()
    raise TypeError(f"Unsupported BC specification: {type(bc)!r}")


# ---------------------------------------------------------------------------
# State generators: conserved Q at interior cell centres, then the periodic
# ghost fill.  `g` is the parameter dict (lower-case names) of the case.
# ---------------------------------------------------------------------------
def _grid(spec):
    from .cases import cell_centres
    x = cell_centres(spec.nx, spec.Lx, spec.order)
    y = cell_centres(spec.ny, spec.Ly, spec.order)
    z = cell_centres(spec.nz, spec.Lz, spec.order)
    return np.meshgrid(z, y, x, indexing="ij")


def _pack(spec, rho, u, v, w, p):
    from .cases import periodic_fill
    Q = np.zeros(spec.state_shape())
    Q[:, :, 0, :] = rho
    Q[:, :, 1, :] = rho * u
    Q[:, :, 2, :] = rho * v
    Q[:, :, 3, :] = rho * w
    Q[:, :, 4, :] = p / (spec.gamma - 1.0) + 0.5 * rho * (u ** 2 + v ** 2 + w ** 2)
    return periodic_fill(Q, spec.order)


def uniform_state(spec, g):
    Z, Y, X = _grid(spec)
    one = np.ones_like(X)
    return _pack(spec, g.get("rho0", 1.0) * one, g.get("u0", 0.0) * one, g.get("v0", 0.0) * one,
                 g.get("w0", 0.0) * one, g.get("p0", 0.71429) * one)


def taylor_green_state(spec, g):
    Z, Y, X = _grid(spec)
    Ma, rho0, p0 = g.get("ma", 0.1), g.get("rho0", 1.0), g.get("p0", 0.71429)
    u = Ma * np.sin(X) * np.cos(Y) * np.cos(Z)
    v = -Ma * np.cos(X) * np.sin(Y) * np.cos(Z)
    p = p0 + rho0 * Ma ** 2 / 16.0 * (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2.0)
    return _pack(spec, rho0 * np.ones_like(X), u, v, 0.0 * X, p)


def riemann_state(spec, g):
    Z, Y, X = _grid(spec)
    left = X < g.get("x0_frac", 0.5) * spec.Lx
    rho = np.where(left, g.get("rho_l", 1.0), g.get("rho_r", 0.125))
    u = np.where(left, g.get("u_l", 0.0), g.get("u_r", 0.0))
    p = np.where(left, g.get("p_l", 1.0), g.get("p_r", 0.1))
    return _pack(spec, rho, u, 0.0 * X, 0.0 * X, p)


def khi_state(spec, g):
    """3D_solver/KHI/set.f90:16-51: stream 2 in 0.25 Ly < y < 0.75 Ly."""
    Z, Y, X = _grid(spec)
    u1, rho1, u2, rho2, p, amp = g["u1"], g["rho1"], g["u2"], g["rho2"], g["p"], g["amp"]
    v = amp * u1 * np.sin(2.0 * np.pi * X / spec.Lx)
    w = amp * u1 * np.sin(2.0 * np.pi * Z / spec.Lz)
    outer = (Y > 0.75 * spec.Lx) | (Y < 0.25 * spec.Lx)
    rho = np.where(outer, rho1, rho2)
    u = np.where(outer, u1, u2)
    return _pack(spec, rho, u, v, w, p * np.ones_like(X))


def nstgv_state(spec, g):
    from .cases import tgv_state
    spec.L0, spec.V0, spec.RHO0, spec.p0, spec.T0 = g["l0"], g["v0"], g["rho0"], g["p0"], g["t"]
    return tgv_state(spec)


def etgv_state(spec, g):
    """3D_solver/ETGV/set.f90:27-40 (non-dimensional TGV, p = 1/gamma + ...)."""
    Z, Y, X = _grid(spec)
    from .cases import periodic_fill
    rho0, M0, gam = g["rho0"], g["m0"], spec.gamma
    Q = np.zeros(spec.state_shape())
    Q[:, :, 0, :] = rho0
    Q[:, :, 1, :] = rho0 * M0 * np.sin(X) * np.cos(Y) * np.cos(Z)
    Q[:, :, 2, :] = -rho0 * M0 * np.cos(X) * np.sin(Y) * np.cos(Z)
    Q[:, :, 4, :] = (1.0 / gam + 0.0625 * rho0 * M0 ** 2 * (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2.0)) / (
        gam - 1.0) + 0.5 * (Q[:, :, 1, :] ** 2 + Q[:, :, 2, :] ** 2) / Q[:, :, 0, :]
    return periodic_fill(Q, spec.order)


def stz_state(spec, g):
    """3D_solver/STZ/set.f90:40-60 on one domain: (rho_L, p_L) below the
    diaphragm at Lz / 2 (the first half of the reference's z slabs), (rho_R,
    p_R) above, at rest, ghost layers included."""
    Z, Y, X = _grid(spec)
    left = Z < 0.5 * spec.Lz
    rho = np.where(left, g.get("rho_l", 1.0), g.get("rho_r", 0.125))
    p = np.where(left, g.get("p_l", 1.0), g.get("p_r", 0.1))
    Q = np.zeros(spec.state_shape())  # set_init writes the ghost cells too: no periodic fill
    Q[:, :, 0, :] = rho
    Q[:, :, 4, :] = p / (spec.gamma - 1.0)
    return Q


CASE_STATES = {"NSTGV": nstgv_state, "ETGV": etgv_state, "KHI": khi_state, "STZ": stz_state}

# Boundary kinds (x, y, z) of the cases whose set_bc is not the periodic fill:
# STZ/set.f90:77-138 (slip walls in x and y, zero-gradient z ends).
CASE_BOUNDARIES = {"STZ": ("slip", "slip", "extrapolate")}


def device_ic(case_name: str, spec, g: dict):
    """The ``ox_ic`` that builds the case's set_init on the device, or None
    where only the host form exists:
      NSTGV (3D_solver/NSTGV/set.f90:16-47): OX_IC_TAYLOR_GREEN with the case
        file's l0, v0, rho0, p0 and t;
      ETGV (3D_solver/ETGV/set.f90:27-40): OX_IC_ETGV with rho0 and m0;
      KHI (3D_solver/KHI/set.f90:31-47): OX_IC_KHI with u1, rho1, u2, rho2,
        p and amp.
    An explicit ic= of the Case, and the other cases, stay on the host."""
    if g.get("ic_type") is not None:
        return None
    from . import _abi
    if case_name == "NSTGV":
        spec.ic = "tgv"
        spec.L0, spec.V0, spec.RHO0, spec.p0, spec.T0 = g["l0"], g["v0"], g["rho0"], g["p0"], g["t"]
        return spec.device_ic()
    if case_name == "ETGV":
        return _abi.OxIc(kind=_abi.OX_IC_ETGV, rho0=float(g["rho0"]), V0=float(g["m0"]))
    if case_name == "KHI":
        return _abi.OxIc(kind=_abi.OX_IC_KHI, V0=float(g["u1"]), rho0=float(g["rho1"]), u2=float(g["u2"]),
                         rho2=float(g["rho2"]), p0=float(g["p"]), amp=float(g["amp"]))
    return None


def initial_state(case_name: str, spec, g: dict) -> np.ndarray:
    ic_type = g.get("ic_type")
    if ic_type == IC_UNIFORM:
        return uniform_state(spec, g)
    if ic_type == IC_TAYLOR_GREEN:
        return taylor_green_state(spec, g)
    if ic_type == IC_RIEMANN:
        return riemann_state(spec, g)
    if case_name in CASE_STATES:
        return CASE_STATES[case_name](spec, g)
    raise ValueError(f"no initial condition for case {case_name!r}; pass ic=UniformIC(...) or similar")
