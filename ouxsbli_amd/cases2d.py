"""Two-dimensional cases: the oblique-shock reflection of 2D_solver/OS.

A Mach-2 free stream enters on the left.  From x = 0.1 Lx on, the top
boundary holds the state behind an oblique shock of angle beta, so the shock
enters the domain there, reflects off the slip wall at the bottom and leaves
on the right (2D_solver/OS/set.f90, OS/mod_globals.f90).  The state is the
physical conserved ``Q(nx,4,ny)`` (numpy shape ``(ny, 4, nx)``), boundary
points included.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import numpy as np

from . import _abi

# OS/mod_globals.f90:5-50
GAMMA = 1.4
R_GAS = 287.03
M0 = 2.0
P_TOT = 100.0e3
T_TOT = 295.0
BETA_REF = math.pi * 37.2 / 180.0
BLT = 1.0e-3
LX = 5.0 * BLT
LY = 2.0 * BLT
NX_REF, NY_REF = 257, 129
DT_REF = 3.0e-9
END_T_REF = 0.1e-3


def free_stream(M: float = M0, gamma: float = GAMMA, R: float = R_GAS, p_tot: float = P_TOT,
                T_tot: float = T_TOT):
    """(u0, p0, T0, rho0) of the isentropic free stream (OS/mod_globals.f90:36-42)."""
    p0 = p_tot / ((1.0 + 0.5 * (gamma - 1.0) * M ** 2) ** (gamma / (gamma - 1.0)))
    T0 = T_tot / (1.0 + 0.5 * (gamma - 1.0) * M ** 2)
    u0 = M * math.sqrt(gamma * R * T0)
    rho0 = p0 / (R * T0)
    return u0, p0, T0, rho0


def deflection(M: float, beta: float, gamma: float = GAMMA) -> float:
    """Flow deflection theta of a shock of angle beta (theta-beta-M relation,
    OS/mod_globals.f90:46)."""
    Ms2 = (M * math.sin(beta)) ** 2
    return math.atan(2.0 * (1.0 / math.tan(beta)) * (Ms2 - 1.0) / (M ** 2 * (gamma + math.cos(2.0 * beta)) + 2.0))


def weak_beta(M: float, theta: float, gamma: float = GAMMA) -> float:
    """Weak-shock angle for deflection theta (radians): the root of the
    theta-beta-M relation between the Mach angle and the angle of maximum
    deflection (the definition of ouxsbli/tests/utils/oblique_shock.py:4-18),
    by bisection."""
    lo = math.asin(1.0 / M) + 1e-12
    hi = 0.5 * math.pi - 1e-12
    # angle of maximum deflection: golden-section search on the unimodal theta(beta)
    a, b = lo, hi
    g = 0.5 * (math.sqrt(5.0) - 1.0)
    for _ in range(200):
        c1, c2 = b - g * (b - a), a + g * (b - a)
        if deflection(M, c1, gamma) > deflection(M, c2, gamma):
            b = c2
        else:
            a = c1
    bmax = 0.5 * (a + b)
    if theta > deflection(M, bmax, gamma):
        raise ValueError("deflection beyond the attached-shock limit")
    a, b = lo, bmax
    for _ in range(200):
        m = 0.5 * (a + b)
        if deflection(M, m, gamma) < theta:
            a = m
        else:
            b = m
    return 0.5 * (a + b)


def post_shock(beta: float, M: float = M0, gamma: float = GAMMA, R: float = R_GAS, p_tot: float = P_TOT,
               T_tot: float = T_TOT):
    """(rho2, ux, uy, p2) behind the incident shock of angle beta, the
    Rankine-Hugoniot state of OS/mod_globals.f90:43-58."""
    u0, p0, T0, rho0 = free_stream(M, gamma, R, p_tot, T_tot)
    Ms = M * math.sin(beta)
    Ms2 = Ms ** 2
    theta = deflection(M, beta, gamma)
    T2 = T0 * (1.0 + 2.0 * (gamma - 1.0) * (Ms2 - 1.0) * (1.0 + gamma * Ms2) / (Ms2 * (gamma + 1.0) ** 2))
    p2 = p0 * (1.0 + 2.0 * gamma * (Ms2 - 1.0) / (gamma + 1.0))
    rho2 = p2 / (R * T2)
    u1 = u0 * math.sin(beta)
    a1 = u0 / M
    u2 = u1 - 2.0 * a1 * (Ms - 1.0 / Ms) / (gamma + 1.0)
    v2 = u0 * math.cos(beta)
    um = math.sqrt(u2 ** 2 + v2 ** 2)
    return rho2, um * math.cos(theta), -um * math.sin(theta), p2


def reflected_state(beta: float, M: float = M0, gamma: float = GAMMA, R: float = R_GAS, p_tot: float = P_TOT,
                    T_tot: float = T_TOT):
    """(rho3, u3, v3 = 0, p3) behind the regular reflection at the slip wall:
    the flow of region 2 (deflected by theta) is turned back by theta
    through a weak shock, so region 3 is parallel to the wall."""
    rho2, ux, uy, p2 = post_shock(beta, M, gamma, R, p_tot, T_tot)
    theta = deflection(M, beta, gamma)
    a2 = math.sqrt(gamma * p2 / rho2)
    V2 = math.hypot(ux, uy)
    M2 = V2 / a2
    br = weak_beta(M2, theta, gamma)
    Mn2 = M2 * math.sin(br)
    p3 = p2 * (1.0 + 2.0 * gamma * (Mn2 ** 2 - 1.0) / (gamma + 1.0))
    rho3 = rho2 * (gamma + 1.0) * Mn2 ** 2 / ((gamma - 1.0) * Mn2 ** 2 + 2.0)
    Mn3 = math.sqrt((1.0 + 0.5 * (gamma - 1.0) * Mn2 ** 2) / (gamma * Mn2 ** 2 - 0.5 * (gamma - 1.0)))
    a3 = math.sqrt(gamma * p3 / rho3)
    V3 = Mn3 * a3 / math.sin(br - theta)
    return rho3, V3, 0.0, p3


def reflected_shock_height(spec: "CaseSpec2D", x: float) -> float:
    """Height of the reflected shock above the wall at abscissa x: the
    incident shock starts on the top boundary at the first post-shock point
    and hits the wall at x_hit; the reflected shock leaves the wall at the
    angle beta_r - theta (beta_r: weak root behind the incident shock)."""
    theta = deflection(spec.M, spec.beta, spec.gamma)
    rho2, ux, uy, p2 = post_shock(spec.beta, spec.M, spec.gamma, spec.R, spec.p_tot, spec.T_tot)
    M2 = math.hypot(ux, uy) / math.sqrt(spec.gamma * p2 / rho2)
    br = weak_beta(M2, theta, spec.gamma)
    x0 = (spec.top_split - 1) * spec.Lx / (spec.nx - 1)
    x_hit = x0 + spec.Ly / math.tan(spec.beta)
    return (x - x_hit) * math.tan(br - theta)


@dataclasses.dataclass
class CaseSpec2D:
    """One 2-D oblique-shock solve: extents incl. the boundary layer, switches,
    constants and the shock angle."""

    nx: int
    ny: int
    order: int = 6
    scheme: str = "SLAU"
    tvd: str = "hybrid"
    slau_variant: str = "HRSLAU2"
    rk: int = 3
    precision: str = "fp64"
    gamma: float = GAMMA
    R: float = R_GAS
    threshold: float = 0.4
    dt: float = DT_REF
    Lx: float = LX
    Ly: float = LY
    device: int = 0
    M: float = M0
    p_tot: float = P_TOT
    T_tot: float = T_TOT
    beta: float = BETA_REF
    seed: Optional[int] = None

    dim = 2

    @property
    def interior_cells(self) -> int:
        return (self.nx - 2) * (self.ny - 2)

    @property
    def stages(self) -> int:
        return 3

    @property
    def top_split(self) -> int:
        return int(0.1 * self.nx)  # OS/set.f90:40,55

    def inflow(self):
        u0, p0, T0, rho0 = free_stream(self.M, self.gamma, self.R, self.p_tot, self.T_tot)
        return rho0, u0, 0.0, p0

    def post(self):
        return post_shock(self.beta, self.M, self.gamma, self.R, self.p_tot, self.T_tot)

    def to_ox(self) -> _abi.Ox2dConfig:
        c = _abi.Ox2dConfig()
        c.abi_version = _abi.OX_ABI_VERSION
        c.nx, c.ny = self.nx, self.ny
        c.order = self.order
        c.scheme = _abi.SCHEME[self.scheme]
        c.tvd = _abi.TVD[self.tvd]
        c.slau_variant = _abi.SLAU_VARIANT[self.slau_variant]
        c.rk = self.rk
        c.precision = _abi.PRECISION[self.precision]
        c.bc = _abi.OX2D_BC_OBLIQUE_SHOCK
        c.top_split = self.top_split
        c.device = self.device
        c.gamma, c.R, c.threshold, c.dt, c.Lx, c.Ly = self.gamma, self.R, self.threshold, self.dt, self.Lx, self.Ly
        for k, v in enumerate(self.inflow()):
            c.inflow[k] = v
        for k, v in enumerate(self.post()):
            c.post[k] = v
        return c

    def state_shape(self):
        """numpy shape of Q(nx,4,ny) in C order (y, var, x)."""
        return (self.ny, 4, self.nx)


def conserved(prim, gamma: float = GAMMA):
    """(rho, rho u, rho v, E) of a primitive (rho, u, v, p), in the operation
    order of OS/set.f90:34-45."""
    rho, u, v, p = prim
    return rho, rho * u, rho * v, p * (1.0 / (gamma - 1.0)) + 0.5 * rho * (u ** 2 + v ** 2)


def os_state(spec: CaseSpec2D) -> np.ndarray:
    """set_init of OS/set.f90:28-46: the free stream everywhere, the
    post-shock state on the top row from i = int(0.1 nx) (1-based) on."""
    Q = np.empty(spec.state_shape(), dtype=np.float64)
    q0 = conserved(spec.inflow(), spec.gamma)
    q2 = conserved(spec.post(), spec.gamma)
    for l in range(4):
        Q[:, l, :] = q0[l]
    s = spec.top_split
    for l in range(4):
        Q[spec.ny - 1, l, s - 1:] = q2[l]
    return Q


def os_spec(nx: int = NX_REF, ny: int = NY_REF, scheme: str = "SLAU", order: int = 6, precision: str = "fp64",
            beta: float = BETA_REF, dt: Optional[float] = None, cfl: Optional[float] = None, **kw) -> CaseSpec2D:
    """The OS case (SLAU ORDER 6, HR-SLAU2, hybrid MUSCL, RK3 on 5 x 2 mm).
    dt: the reference's 3e-9 s at 257 x 129 unless given; ``cfl`` instead sets
    dt = cfl * min(dx, dy) / (u0 + c0)."""
    spec = CaseSpec2D(nx=nx, ny=ny, order=order, scheme=scheme, precision=precision, beta=beta, **kw)
    if cfl is not None:
        u0, p0, T0, rho0 = free_stream(spec.M, spec.gamma, spec.R, spec.p_tot, spec.T_tot)
        c0 = math.sqrt(spec.gamma * p0 / rho0)
        h = min(spec.Lx / (nx - 1), spec.Ly / (ny - 1))
        spec.dt = cfl * h / (u0 + c0)
    elif dt is not None:
        spec.dt = dt
    return spec


# Courant number of the reference OS run (dt 3e-9 s at 257 x 129), used to
# size dt on the other grids of the batched ensemble.
CFL_REF = DT_REF * (free_stream()[0] + math.sqrt(GAMMA * R_GAS * free_stream()[2])) / min(LX / (NX_REF - 1),
                                                                                      LY / (NY_REF - 1))


def oblique_shock_specs(ncase: int = 64, seed: int = 20261206, precision: str = "fp32_bulk", scheme: str = "Roe",
                  order: int = 6, nx_range=(128, 512), ny_range=(64, 256), theta_deg=(4.0, 12.0)) -> List[CaseSpec2D]:
    """The batched oblique-shock ensemble: interior sizes nx-2 ~ U{128..512}
    and ny-2 ~ U{64..256} drawn independently, deflection theta ~ U(4, 12)
    degrees, beta the weak root at Mach 2; Roe + MUSCL (ORDER 6, hybrid
    limiter), RK3, dt at the reference run's Courant number."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(ncase):
        ix = int(rng.integers(nx_range[0], nx_range[1] + 1))
        iy = int(rng.integers(ny_range[0], ny_range[1] + 1))
        th = math.radians(float(rng.uniform(theta_deg[0], theta_deg[1])))
        beta = weak_beta(M0, th)
        out.append(os_spec(ix + 2, iy + 2, scheme=scheme, order=order, precision=precision, beta=beta,
                           cfl=CFL_REF, seed=seed + 1 + k))
    return out
