"""Case specifications: grids, physical constants and initial conditions.

The reference encodes each case as a Fortran ``module set`` plus
``mod_globals.f90`` parameters and builds one executable per case.  Here a case
is a :class:`CaseSpec` (the ``ox_config`` switches plus the physical constants
the initial condition needs) and a numpy initial-condition generator that
returns the physical conserved state ``Q(nx,5,ny,nz)`` (Fortran order, ghosts
filled periodically) ready for ``ox_set_state``.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
from typing import Optional

import numpy as np

from . import _abi

# Thermodynamic constants shared by the 3D cases (3D_solver/NSTGV/mod_globals.f90:39-43).
GAMMA = 1.4
R_GAS = 287.03
PR = 0.71
PRT = 0.9
T_NSTGV = 530.0 * 5.0 / 9.0  # NSTGV/mod_globals.f90:47
SUTHERLAND_S = 111.0


def sutherland_mu(T: float) -> float:
    """mu0 of NSTGV/mod_globals.f90:49 (Sutherland's law, S = 111 K)."""
    return 1.716e-5 * (273.2 + SUTHERLAND_S) / (T + SUTHERLAND_S) * (T / 273.2) ** 1.5


@dataclasses.dataclass
class CaseSpec:
    """One solve: extents incl. ghosts, scheme switches and constants."""

    nx: int
    ny: int
    nz: int
    order: int = 4
    scheme: str = "KEEP"
    visc: str = "NS"
    visc_order: int = 0
    visc_stencil: str = "Central"
    tvd: str = "hybrid"
    slau_variant: str = "HRSLAU2"
    rk: int = 4
    precision: str = "fp64"
    gamma: float = GAMMA
    R: float = R_GAS
    Pr: float = PR
    Prt: float = PRT
    threshold: float = 0.4
    dt: float = 1e-3
    Lx: float = 2.0 * math.pi
    Ly: float = 2.0 * math.pi
    Lz: float = 2.0 * math.pi
    device: int = 0
    # boundary kind of x, y, z (both ends): "periodic", "extrapolate" (zero
    # gradient) or "slip" (slip wall)
    bc: tuple = ("periodic", "periodic", "periodic")
    # initial-condition constants (not part of ox_config)
    ic: str = "tgv"
    L0: float = 1.0
    V0: float = 1.0
    RHO0: float = 1.0
    p0: float = 1.0
    T0: float = T_NSTGV
    seed: Optional[int] = None

    @property
    def ghost(self) -> int:
        return self.order // 2

    @property
    def interior(self):
        g = self.ghost
        return (self.nx - 2 * g, self.ny - 2 * g, self.nz - 2 * g)

    @property
    def interior_cells(self) -> int:
        a, b, c = self.interior
        return a * b * c

    @property
    def stages(self) -> int:
        return 3 if self.rk == 3 else 4

    def to_ox(self) -> _abi.OxConfig:
        c = _abi.OxConfig()
        c.abi_version = _abi.OX_ABI_VERSION
        c.nx, c.ny, c.nz = self.nx, self.ny, self.nz
        c.order = self.order
        c.scheme = _abi.SCHEME[self.scheme]
        c.visc = _abi.VISC[self.visc]
        c.visc_order = self.visc_order
        c.visc_stencil = _abi.STENCIL[self.visc_stencil]
        c.tvd = _abi.TVD[self.tvd]
        c.slau_variant = _abi.SLAU_VARIANT[self.slau_variant]
        c.rk = self.rk
        c.precision = _abi.PRECISION[self.precision]
        c.bc_x, c.bc_y, c.bc_z = (_abi.BC[b] for b in self.bc)
        c.device = self.device
        c.gamma, c.R, c.Pr, c.Prt = self.gamma, self.R, self.Pr, self.Prt
        c.threshold = self.threshold
        c.dt = self.dt
        c.Lx, c.Ly, c.Lz = self.Lx, self.Ly, self.Lz
        return c

    def device_ic(self):
        """The ``ox_ic`` that builds this case's initial state on the device
        (ox_init_state), or None where only the host builds it."""
        if self.ic == "tgv":
            return _abi.OxIc(kind=_abi.OX_IC_TAYLOR_GREEN, L0=self.L0, V0=self.V0, rho0=self.RHO0, p0=self.p0,
                             T0=self.T0)
        if self.ic == "khi_tanh":
            return _abi.OxIc(kind=_abi.OX_IC_SHEAR_LAYER, seed=int(self.seed), V0=self.V0, rho0=self.RHO0,
                             p0=self.p0, delta=self.Ly / 28.0, noise=1e-3 * self.V0)
        if self.ic == "dhit":
            n = self.nx - self.order
            H = dhit_modes(n, int(self.seed))
            ic = _abi.OxIc(kind=_abi.OX_IC_SPECTRAL, kmax=H.shape[1] - 1, seed=int(self.seed), V0=self.V0,
                           rho0=self.RHO0, p0=self.p0,
                           modes=H.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
            ic._modes = H  # the host array lives as long as the descriptor
            return ic
        return None

    def state_shape(self):
        """numpy shape of Q(nx,5,ny,nz) in C order (z, y, var, x)."""
        return (self.nz, self.ny, 5, self.nx)


def cell_centres(n: int, L: float, order: int) -> np.ndarray:
    """Cell centres xc(1..n) of set_grid_cyclic{2,4,6}_3D (src/set_coordinate.f90:120-339).

    Node i sits at dx1*(i-1-order/2) with dx1 = L/(n-order); ghost nodes extend
    the same spacing; centres are midpoints.
    """
    off = order // 2
    dx1 = L / float(n - order)
    idx = np.arange(1, n + 2, dtype=np.float64)
    x = dx1 * (idx - 1 - off)
    return 0.5 * (x[:-1] + x[1:])


def periodic_fill(Q: np.ndarray, order: int) -> np.ndarray:
    """Ghost layers = periodic image of the interior (set_bc_cyclic*_init,
    3D_solver/src/set_bc_common.f90:11-97).  Q has shape (nz, ny, 5, nx)."""
    h = order // 2
    nz, ny, _, nx = Q.shape
    ix, iy, iz = nx - 2 * h, ny - 2 * h, nz - 2 * h
    Q[:, :, :, :h] = Q[:, :, :, ix:ix + h]
    Q[:, :, :, nx - h:] = Q[:, :, :, h:2 * h]
    Q[:, :h, :, :] = Q[:, iy:iy + h, :, :]
    Q[:, ny - h:, :, :] = Q[:, h:2 * h, :, :]
    Q[:h, :, :, :] = Q[iz:iz + h, :, :, :]
    Q[nz - h:, :, :, :] = Q[h:2 * h, :, :, :]
    return Q


def tgv_spec(n: int, order: int, Ma: float, Re: float = 1600.0, cfl: float = 0.3, L0: float = 1.0,
             cfl_mode: str = "reference", **kw) -> CaseSpec:
    """Taylor-Green vortex on [0, 2*pi*L0)^3 (3D_solver/NSTGV/mod_globals.f90:10-58).

    ``n`` is the interior size per side; extents add ``order/2`` ghosts per side.
    cfl_mode "reference": dt = CFL * (Lx/(nx-1)) / V0, the NSTGV formula
    (mod_globals.f90:54; nx counts ghosts), a convective Courant number.
    cfl_mode "acoustic": dt = CFL * dx / (V0 + c0) with dx = Lx/n, the Courant
    number of the fastest wave.  At Ma 0.1 the reference formula with CFL 0.3
    puts the acoustic Courant number at 3.2, beyond RK4's stability limit.
    """
    N = n + order
    T0 = T_NSTGV
    mu0 = sutherland_mu(T0)
    V0 = Ma * math.sqrt(GAMMA * R_GAS * T0)
    RHO0 = mu0 * Re / (V0 * L0)
    p0 = RHO0 * R_GAS * T0
    L = 2.0 * math.pi * L0
    if cfl_mode == "reference":
        dt = cfl * (L / float(N - 1)) / V0
    elif cfl_mode == "acoustic":
        c0 = math.sqrt(GAMMA * R_GAS * T0)
        dt = cfl * (L / float(n)) / (V0 + c0)
    else:
        raise ValueError(cfl_mode)
    spec = CaseSpec(nx=N, ny=N, nz=N, order=order, dt=dt, Lx=L, Ly=L, Lz=L, ic="tgv", L0=L0, V0=V0,
                    RHO0=RHO0, p0=p0, T0=T0)
    for k, v in kw.items():
        setattr(spec, k, v)
    return spec


def tgv_state(spec: CaseSpec) -> np.ndarray:
    """NSTGV set_init (3D_solver/NSTGV/set.f90:16-47): conserved Q at interior
    cell centres, then the periodic ghost fill.  Returns shape (nz, ny, 5, nx)."""
    L0, V0, RHO0, p0, T = spec.L0, spec.V0, spec.RHO0, spec.p0, spec.T0
    x = cell_centres(spec.nx, spec.Lx, spec.order)
    y = cell_centres(spec.ny, spec.Ly, spec.order)
    z = cell_centres(spec.nz, spec.Lz, spec.order)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    p = p0 + RHO0 * (V0 ** 2) * (np.cos(2.0 * X / L0) + np.cos(2.0 * Y / L0)) * (np.cos(2.0 * Z / L0) + 2.0) / 16.0
    rho = p / (spec.R * T)
    Q = np.zeros(spec.state_shape(), dtype=np.float64)
    Q[:, :, 0, :] = rho
    Q[:, :, 1, :] = rho * V0 * np.sin(X / L0) * np.cos(Y / L0) * np.cos(Z / L0)
    Q[:, :, 2, :] = -rho * V0 * np.cos(X / L0) * np.sin(Y / L0) * np.cos(Z / L0)
    Q[:, :, 3, :] = 0.0
    Q[:, :, 4, :] = p / (spec.gamma - 1.0) + 0.5 * (Q[:, :, 1, :] ** 2 + Q[:, :, 2, :] ** 2) / Q[:, :, 0, :]
    return periodic_fill(Q, spec.order)


# ---------------------------------------------------------------------------
# Shock tube along z between slip walls (the STZ case family of
# 3D_solver/STZ: Euler SLAU (variant SLAU) ORDER 2, TVD, RK3; slip walls in x and y,
# zero-gradient ends in z) on one domain instead of MPI z-slabs.
# ---------------------------------------------------------------------------
def stz_spec(nz: int = 256, nt: int = 4, order: int = 2, scheme: str = "SLAU", precision: str = "fp64",
             rk: int = 3, cfl: float = 0.01) -> CaseSpec:
    """STZ/mod_globals.f90: Lx = Ly = 0.1, Lz = 1, (rho, p) = (1, 1) | (0.125,
    0.1), gamma 1.4, dt = CFL Lz / ((Nz - 1) sqrt(p_L / rho_L)) with Nz the
    z extent; STZ/config.fypp: SLAU variant SLAU, TVD, ORDER 2, RK3; STZ/set.f90:77-138
    walls.  `nz` and `nt` are the interior cells along z and across."""
    Nz = nz + order
    dt = cfl * 1.0 / ((Nz - 1) * math.sqrt(1.0 / 1.0))
    return CaseSpec(nx=nt + order, ny=nt + order, nz=Nz, order=order, scheme=scheme, visc="Euler", tvd="tvd",
                    slau_variant="SLAU", rk=rk, precision=precision, gamma=1.4, R=287.03, threshold=0.4, dt=dt,
                    Lx=0.1, Ly=0.1, Lz=1.0, bc=("slip", "slip", "extrapolate"), ic="stz", RHO0=1.0, p0=1.0)


STZ_LEFT = (1.0, 1.0)      # (rho_L, p_L), STZ/mod_globals.f90
STZ_RIGHT = (0.125, 0.1)   # (rho_R, p_R)


def stz_state(spec: CaseSpec) -> np.ndarray:
    """STZ/set.f90:40-60 on one domain: the left state where the cell centre
    is below Lz / 2, the right state above, at rest; ghost layers included."""
    h = spec.order // 2
    dz = spec.Lz / (spec.nz - spec.order)
    zc = (np.arange(spec.nz) - h + 0.5) * dz
    left = zc < 0.5 * spec.Lz
    Q = np.zeros(spec.state_shape())
    Q[:, :, 0, :] = np.where(left, STZ_LEFT[0], STZ_RIGHT[0])[:, None, None]
    Q[:, :, 4, :] = (np.where(left, STZ_LEFT[1], STZ_RIGHT[1]) / (spec.gamma - 1.0))[:, None, None]
    return Q


def initial_state(spec: CaseSpec) -> np.ndarray:
    if spec.ic == "tgv":
        return tgv_state(spec)
    if spec.ic == "stz":
        return stz_state(spec)
    if spec.ic == "khi_tanh":
        return khi_state(spec)
    if spec.ic == "dhit":
        return dhit_state(spec)
    raise ValueError(f"unknown initial condition {spec.ic!r}")


# The benchmark workloads (interior sizes; extents add the ghosts).
def tgv_keep4_spec(n: int = 64) -> CaseSpec:
    """Plain FP64 NS TGV, KEEP ORDER 4, central-4 viscous, RK4, Ma 0.1, acoustic CFL 0.3."""
    return tgv_spec(n, 4, Ma=0.1, cfl=0.3, cfl_mode="acoustic", scheme="KEEP", visc="NS", rk=4, tvd="none")


def supersonic_tgv_spec(n: int = 256, precision: str = "fp64") -> CaseSpec:
    """Supersonic NS TGV, Ma 1.25, SLAU ORDER 6 HR-SLAU2, hybrid TVD (threshold 0.4),
    central-6 viscous, RK3-TVD, CFL 0.03 (the NSTGV case file)."""
    return tgv_spec(n, 6, Ma=1.25, cfl=0.03, scheme="SLAU", visc="NS", rk=3, tvd="hybrid",
                    slau_variant="HRSLAU2", threshold=0.4, precision=precision)


# ---------------------------------------------------------------------------
# Serving workload: 3-D Kelvin-Helmholtz solves (the KHI case family of
# 3D_solver/KHI: Euler, RK3, no TVD, HR-SLAU2 for SLAU), here with a periodic
# double tanh shear layer of thickness L/28 at Mach 0.5 and seeded
# velocity noise of 1e-3 U, on [0, 1)^3 in non-dimensional units
# (rho0 = 1, p0 = 1/gamma, so c0 = 1).
# ---------------------------------------------------------------------------
def khi_spec(n: int, order: int, scheme: str, precision: str = "fp64", seed: int = 0, Ma: float = 0.5,
             cfl: float = 0.3) -> CaseSpec:
    N = n + order
    L = 1.0
    c0 = 1.0
    U = Ma * c0
    dt = cfl * (L / n) / (U + c0)
    spec = CaseSpec(nx=N, ny=N, nz=N, order=order, scheme=scheme, visc="Euler", visc_order=0, tvd="none",
                    slau_variant="HRSLAU2", rk=3, precision=precision, dt=dt, Lx=L, Ly=L, Lz=L, ic="khi_tanh",
                    V0=U, RHO0=1.0, p0=1.0 / GAMMA, seed=seed)
    return spec


_SM_GAMMA = 0x9E3779B97F4A7C15
_U64 = (1 << 64) - 1


def splitmix64(z: np.ndarray) -> np.ndarray:
    """SplitMix64's output function (Steele, Lea and Flood, OOPSLA 2014) on a
    uint64 array: the counter plus the golden-ratio increment, then two
    xor-shift-multiply rounds and a final xor-shift, all modulo 2^64."""
    with np.errstate(over="ignore"):
        z = z + np.uint64(_SM_GAMMA)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def cell_noise(seed: int, shape, field: int) -> np.ndarray:
    """Uniform noise in [-1, 1) of velocity component `field` (0, 1, 2) at
    every cell of the padded (nz, ny, nx) array: SplitMix64 of the key
    (seed + 1) * 0x9E3779B97F4A7C15 + 3 m + field for linear cell m, whose top
    53 bits are scaled to [0, 1) exactly.  The device kernel of
    ox_init_state (ouxsbli_amd/csrc/ox_ic.hip) draws the same numbers."""
    base = ((int(seed) + 1) * _SM_GAMMA) & _U64
    m = np.arange(int(np.prod(shape)), dtype=np.uint64)
    with np.errstate(over="ignore"):
        key = np.uint64(base) + (np.uint64(3) * m + np.uint64(field))
    u01 = (splitmix64(key) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return (2.0 * u01 - 1.0).reshape(shape)


def khi_state(spec: CaseSpec) -> np.ndarray:
    """Double tanh shear layer u(y) = U/2 * (tanh((y-1/4)/d) - tanh((y-3/4)/d) - 1),
    d = L/28, plus seeded noise of amplitude 1e-3 U in u, v, w (uniform in
    [-1, 1), a counter-based draw per cell and component, see cell_noise);
    uniform density and pressure.  The host form of the OX_IC_SHEAR_LAYER
    initial condition the library builds on the device."""
    y = cell_centres(spec.ny, spec.Ly, spec.order)
    U, d = spec.V0, spec.Ly / 28.0
    shape = (spec.nz, spec.ny, spec.nx)
    prof = 0.5 * U * (np.tanh((y - 0.25 * spec.Ly) / d) - np.tanh((y - 0.75 * spec.Ly) / d) - 1.0)
    amp = 1e-3 * U
    u = cell_noise(spec.seed, shape, 0)
    u *= amp
    u += prof[None, :, None]
    v = cell_noise(spec.seed, shape, 1)
    v *= amp
    w = cell_noise(spec.seed, shape, 2)
    w *= amp
    rho, p = spec.RHO0, spec.p0
    Q = np.empty(spec.state_shape(), dtype=np.float64)
    Q[:, :, 0, :] = rho
    Q[:, :, 1, :] = rho * u
    Q[:, :, 2, :] = rho * v
    Q[:, :, 3, :] = rho * w
    ke = u * u
    ke += v * v
    ke += w * w
    Q[:, :, 4, :] = p / (spec.gamma - 1.0) + 0.5 * rho * ke
    return periodic_fill(Q, spec.order)


# ---------------------------------------------------------------------------
# Decaying isotropic turbulence (LES): a random-phase solenoidal velocity field
# built mode by mode in the two-vector frame normal to k (the construction of
# 3D_solver/DHIT/set_init_dhit.f90:96-169, Rogallo 1981), here with the
# spectrum E(k) ~ k^4 exp(-2 (k/k0)^2), k0 = 4, on the half spectrum of a real
# transform, and scaled to the turbulent Mach number Mt = sqrt(<|u|^2>)/c0 =
# u' sqrt(3)/c0.  Uniform density and temperature in physical units (the
# viscosity follows Sutherland's law).  The three random phases of a mode are
# a counter-based draw keyed by (seed, kx, ky, kz), so a mode's coefficients
# do not depend on the grid: the host form inverts the whole half spectrum
# with an inverse real FFT, the library (ox_init_state, OX_IC_SPECTRAL)
# synthesises the field on the device from the modes with |k| <= DHIT_KMAX,
# beyond which every amplitude is below 1e-20 of the largest.
# ---------------------------------------------------------------------------
DHIT_K0 = 4.0
DHIT_KMAX = 28


def dhit_spec(n: int, precision: str = "fp32_bulk", seed: int = 20261203, Mt: float = 0.3,
              cfl: float = 0.3, order: int = 6) -> CaseSpec:
    N = n + order
    L = 2.0 * math.pi
    T0, RHO0 = 300.0, 1.2
    c0 = math.sqrt(GAMMA * R_GAS * T0)
    urms = Mt * c0  # sqrt(<|u|^2>)
    dt = cfl * (L / n) / (2.0 * urms + c0)
    return CaseSpec(nx=N, ny=N, nz=N, order=order, scheme="Hybrid", visc="LES", visc_order=6 if order == 6 else 4,
                    tvd="hybrid", slau_variant="HRSLAU2", threshold=0.4, rk=4, precision=precision, dt=dt,
                    Lx=L, Ly=L, Lz=L, ic="dhit", V0=urms, RHO0=RHO0, T0=T0, p0=RHO0 * R_GAS * T0, seed=seed)


def _mode_phases(seed: int, KX, KY, KZ):
    """The three phases in [0, 2 pi) of each mode (integer wavenumbers, kx >= 0):
    SplitMix64 of (seed + 1) * 0x9E3779B97F4A7C15 + 3 id + c with the mode id
    ((kz + 2^20) 2^21 + (ky + 2^20)) 2^21 + kx, top 53 bits scaled to [0, 1)."""
    base = ((int(seed) + 1) * _SM_GAMMA) & _U64
    off = np.int64(1 << 20)
    mid = ((KZ.astype(np.int64) + off) * np.int64(1 << 21) + (KY.astype(np.int64) + off)) * np.int64(1 << 21) + \
        KX.astype(np.int64)
    mid = mid.astype(np.uint64)
    out = []
    with np.errstate(over="ignore"):
        for c in range(3):
            key = np.uint64(base) + (np.uint64(3) * mid + np.uint64(c))
            out.append((splitmix64(key) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53 * (2.0 * np.pi))
    return out


def _dhit_coeffs(KX, KY, KZ, n: int, seed: int, k0: float = DHIT_K0):
    """Spectral coefficients (u_hat, v_hat, w_hat) of the modes on the given
    integer wavenumber arrays (kx >= 0); zero at k = 0 and at the Nyquist
    wavenumbers of an n^3 grid."""
    kmag = np.sqrt(KX * KX + KY * KY + KZ * KZ)
    k12 = np.sqrt(KX * KX + KY * KY)
    nyq = (np.abs(KX) == n // 2) | (np.abs(KY) == n // 2) | (np.abs(KZ) == n // 2)
    ph = _mode_phases(seed, KX, KY, KZ)
    with np.errstate(divide="ignore", invalid="ignore"):
        E = kmag ** 4 * np.exp(-2.0 * (kmag / k0) ** 2)
        amp = np.where((kmag > 0) & ~nyq, np.sqrt(E / (2.0 * np.pi * kmag * kmag)), 0.0)
        a = amp * np.exp(1j * ph[0]) * np.cos(ph[2])
        b = amp * np.exp(1j * ph[1]) * np.sin(ph[2])
        inplane = k12 > 0
        s12 = np.where(inplane, k12, 1.0)
        sk = np.where(kmag > 0, kmag, 1.0)
        ux = np.where(inplane, (KY / s12) * a + (KX / s12) * (KZ / sk) * b, a)
        uy = np.where(inplane, -(KX / s12) * a + (KY / s12) * (KZ / sk) * b, 0.0)
        uz = np.where(inplane, -(k12 / sk) * b, 0.0)
    return ux, uy, uz


def dhit_velocity(n: int, seed: int, k0: float = DHIT_K0):
    """Solenoidal random-phase velocity (3, n, n, n) in (z, y, x) order with
    E(k) ~ k^4 exp(-2 (k/k0)^2) (integer wavenumbers of the 2*pi box), not
    yet scaled: the inverse real FFT of the whole half spectrum."""
    k1 = np.fft.fftfreq(n, d=1.0 / n)
    kr = np.fft.rfftfreq(n, d=1.0 / n)
    KZ, KY, KX = np.meshgrid(k1, k1, kr, indexing="ij")
    ux, uy, uz = _dhit_coeffs(KX, KY, KZ, n, seed, k0)
    try:
        import scipy.fft as sfft
        inv = lambda c: sfft.irfftn(c, s=(n, n, n), workers=-1)
    except ImportError:
        inv = lambda c: np.fft.irfftn(c, s=(n, n, n))
    return np.stack([inv(c) for c in (ux, uy, uz)])


def dhit_modes(n: int, seed: int, kmax: int = DHIT_KMAX, k0: float = DHIT_K0) -> np.ndarray:
    """The modes the device synthesis sums: complex array (3, kx, ky, kz)
    over kx = 0..K, ky, kz = -K..K with K = min(kmax, n/2 - 1), zero outside
    the ball |k| <= kmax; the same coefficients dhit_velocity inverts."""
    K = min(kmax, n // 2 - 1)
    kx = np.arange(0, K + 1, dtype=np.float64)
    ks = np.arange(-K, K + 1, dtype=np.float64)
    KX, KY, KZ = np.meshgrid(kx, ks, ks, indexing="ij")
    ux, uy, uz = _dhit_coeffs(KX, KY, KZ, n, seed, k0)
    H = np.stack([ux, uy, uz]).astype(np.complex128)
    H[:, (KX * KX + KY * KY + KZ * KZ) > kmax * kmax] = 0.0
    return np.ascontiguousarray(H)


def dhit_state(spec: CaseSpec) -> np.ndarray:
    n = spec.nx - spec.order
    u = dhit_velocity(n, spec.seed)
    u *= spec.V0 / math.sqrt(float(np.mean(np.sum(u * u, axis=0))))
    h = spec.order // 2
    Q = np.zeros(spec.state_shape(), dtype=np.float64)
    rho, p = spec.RHO0, spec.p0
    sl = (slice(h, h + n), slice(h, h + n))
    Q[sl[0], sl[1], 0, h:h + n] = rho
    Q[sl[0], sl[1], 1, h:h + n] = rho * u[0]
    Q[sl[0], sl[1], 2, h:h + n] = rho * u[1]
    Q[sl[0], sl[1], 3, h:h + n] = rho * u[2]
    Q[sl[0], sl[1], 4, h:h + n] = p / (spec.gamma - 1.0) + 0.5 * rho * np.sum(u * u, axis=0)
    return periodic_fill(Q, spec.order)
