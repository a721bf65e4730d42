"""Non-periodic directions (slip walls, zero-gradient ends) on the oracle.

The reference's non-periodic cases fill their ghost layers with a case
set_bc after every stage; 3D_solver/STZ/set.f90:77-138 holds the two kinds
built here: slip walls (ghost ovlp-m mirrors ovlp+1+m, the normal momentum
negated) and zero-gradient ends (every ghost layer copies the adjacent
interior cell).  Pins:

* the oracle's fill against a literal restatement of those Fortran loops;
* a slip wall is a mirror: next to slip walls the right-hand side equals, bit
  for bit, that of the periodic case of twice the length whose other half is
  the mirror image (reversed, normal momentum negated), along each axis, for
  schemes that read the Ducros sensor of the ghost layer (Hybrid, SLAU with
  MUSCL at ORDER 6) and for KEEP (over whole steps the periodic run need not
  keep its two halves mirror images bit for bit, so the comparison is per
  right-hand side, on the initial field and on an evolved one);
* Sod's problem along z between slip walls (the STZ case on one domain)
  against the exact solution.
"""
import pathlib
import sys

import numpy as np
import pytest

from ouxsbli_amd import cases
from ouxsbli_amd.cases import CaseSpec
from oracle_lib import Oracle

GOLD = pathlib.Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(GOLD))
import make_sod  # noqa: E402


def _stz_fill_literal(QJ, h):
    """3D_solver/STZ/set.f90:77-138 with Fortran indices shifted by one
    (QJ has numpy shape (nz, ny, 5, nx)): x slip over all (j, k), y slip over
    all (i, k), then the two zero-gradient z ends over all (i, j)."""
    Q = QJ.copy()
    nz, ny, _, nx = Q.shape
    ovlp = h
    for m in range(ovlp):
        for l in range(1, 6):
            s = -1.0 if l == 2 else 1.0
            Q[:, :, l - 1, (ovlp - m) - 1] = s * Q[:, :, l - 1, (ovlp + 1 + m) - 1]
            Q[:, :, l - 1, (nx - ovlp + 1 + m) - 1] = s * Q[:, :, l - 1, (nx - ovlp - m) - 1]
    for m in range(ovlp):
        for l in range(1, 6):
            s = -1.0 if l == 3 else 1.0
            Q[:, (ovlp - m) - 1, l - 1, :] = s * Q[:, (ovlp + 1 + m) - 1, l - 1, :]
            Q[:, (ny - ovlp + 1 + m) - 1, l - 1, :] = s * Q[:, (ny - ovlp - m) - 1, l - 1, :]
    for k in range(1, ovlp + 1):
        Q[k - 1] = Q[(ovlp + 1) - 1]
    for k in range(1, ovlp + 1):
        Q[(nz - ovlp + k) - 1] = Q[(nz - ovlp) - 1]
    return Q


@pytest.mark.parametrize("order", [2, 4, 6])
def test_fill_restates_stz_set_bc(order):
    spec = cases.stz_spec(nz=12, nt=8, order=order)
    rng = np.random.default_rng(order)
    QJ = 1.0 + rng.random(spec.state_shape())
    o = Oracle(spec)
    o.set_qj(QJ)
    o.set_bc()
    np.testing.assert_array_equal(o.get_qj(), _stz_fill_literal(QJ, order // 2))


def test_periodic_direction_between_walls():
    """A periodic direction among non-periodic ones takes its periodic image
    (set_bc_cyclic), the walls their mirror, in any corner."""
    spec = cases.stz_spec(nz=12, nt=6, order=4)
    spec.bc = ("periodic", "slip", "extrapolate")
    rng = np.random.default_rng(3)
    QJ = 1.0 + rng.random(spec.state_shape())
    o = Oracle(spec)
    o.set_qj(QJ)
    o.set_bc()
    Q = o.get_qj()
    h, n = 2, spec.nx - 4
    np.testing.assert_array_equal(Q[..., :h], Q[..., n:n + h])
    np.testing.assert_array_equal(Q[..., n + h:], Q[..., h:2 * h])
    # y slip mirror, with the x images in place
    np.testing.assert_array_equal(Q[:, h - 1, [0, 1, 3, 4], :], Q[:, h, [0, 1, 3, 4], :])
    np.testing.assert_array_equal(Q[:, h - 1, 2, :], -Q[:, h, 2, :])


def _field(shape_zyx, seed):
    """A smooth random field with one oblique jump (the sensor switches on
    near it), physical conserved variables (nz, ny, 5, nx)."""
    nz, ny, nx = shape_zyx
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    rho = 1.0 + 0.1 * rng.random((nz, ny, nx)) + 0.5 * ((x + 0.7 * y + 0.4 * z) > 0.5 * (nx + ny + nz) / 2.2)
    u = 0.3 * np.sin(0.7 * x + 0.3 * y) + 0.05 * rng.standard_normal((nz, ny, nx))
    v = 0.2 * np.cos(0.5 * y + 0.4 * z) + 0.05 * rng.standard_normal((nz, ny, nx))
    w = 0.25 * np.sin(0.6 * z + 0.2 * x) + 0.05 * rng.standard_normal((nz, ny, nx))
    p = 1.0 + 0.1 * rng.random((nz, ny, nx)) + 0.8 * ((x + 0.7 * y + 0.4 * z) > 0.5 * (nx + ny + nz) / 2.2)
    Q = np.empty((nz, ny, 5, nx))
    Q[:, :, 0] = rho
    Q[:, :, 1] = rho * u
    Q[:, :, 2] = rho * v
    Q[:, :, 3] = rho * w
    Q[:, :, 4] = p / 0.4 + 0.5 * rho * (u * u + v * v + w * w)
    return Q


MIRROR_SCHEMES = {
    # (order, scheme, tvd, slau_variant, rk[, visc, visc_order, visc_stencil])
    "hybrid6": (6, "Hybrid", "hybrid", "HRSLAU2", 3),
    "slau6-muscl": (6, "SLAU", "tvd", "HRSLAU2", 3),
    "keep4": (4, "KEEP", "none", "HRSLAU2", 4),
    # viscous: the wall face's stress reads the cross derivatives of the
    # ghost cells, formed from the filled primitives
    "slau6-ns6": (6, "SLAU", "hybrid", "HRSLAU2", 3, "NS", 6, "Central"),
    "hybrid6-me4": (6, "Hybrid", "hybrid", "HRSLAU2", 3, "NS", 6, "ME4Base"),
    "keep4-ns4": (4, "KEEP", "none", "HRSLAU2", 4, "NS", 4, "Central"),
}
MIRROR_INTERIOR = (14, 10, 12)  # (nz, ny, nx) of the case between walls


def mirror_pair(name, axis, precision="fp64"):
    """(spec_wall, Q_wall, spec_periodic, Q_periodic, half): the case between
    slip walls along `axis` (0 x, 1 y, 2 z; periodic across), and the periodic
    case of twice its length whose first half is its mirror image; `half`
    selects the second half of the periodic case's interior."""
    order, scheme, tvd, variant, rk, *visc = MIRROR_SCHEMES[name]
    visc, vorder, stencil = visc if visc else ("Euler", 0, "Central")
    h = order // 2
    nz, ny, nx = MIRROR_INTERIOR
    L = {"x": 1.0, "y": 0.8, "z": 1.2}
    bc = ["periodic"] * 3
    bc[axis] = "slip"
    common = dict(order=order, scheme=scheme, visc=visc, visc_order=vorder, visc_stencil=stencil, tvd=tvd,
                  slau_variant=variant, rk=rk, precision=precision, threshold=0.4, dt=2e-3, T0=300.0)
    wall = CaseSpec(nx=nx + order, ny=ny + order, nz=nz + order, Lx=L["x"], Ly=L["y"], Lz=L["z"], bc=tuple(bc),
                    **common)
    ext = [nx, ny, nz]
    ext[axis] *= 2
    Ls = [L["x"], L["y"], L["z"]]
    Ls[axis] *= 2
    per = CaseSpec(nx=ext[0] + order, ny=ext[1] + order, nz=ext[2] + order, Lx=Ls[0], Ly=Ls[1], Lz=Ls[2], **common)
    f = _field((nz, ny, nx), seed=10 + axis)
    npax = {0: 3, 1: 1, 2: 0}[axis]  # numpy axis of x, y, z in (nz, ny, 5, nx)
    mir = np.flip(f, axis=npax).copy()
    mir[:, :, 1 + axis] *= -1.0
    g = np.concatenate([mir, f], axis=npax)
    Qp = np.zeros(per.state_shape())
    Qp[h:-h, h:-h, :, h:-h] = g
    Qp = cases.periodic_fill(Qp, order)
    # the case between walls: its interior, ghosts by its own fill
    Qw = np.zeros(wall.state_shape())
    Qw[h:-h, h:-h, :, h:-h] = f
    o = Oracle(wall)
    o.set_qj(Qw)  # Jacobian 1 is not assumed: the fill acts on QJ alike
    o.set_bc()
    Qw = o.get_qj()
    del o
    sl = [slice(h, -h), slice(h, -h), slice(None), slice(h, -h)]
    sl[npax] = slice(h + [nx, ny, nz][axis], -h)
    return wall, Qw, per, Qp, tuple(sl)


def _interior(spec):
    h = spec.order // 2
    return (slice(h, -h), slice(h, -h), slice(None), slice(h, -h))


def _rhs_pair(wall, Qw, per, Qp, half):
    a = Oracle(wall)
    a.set_qj(Qw)
    b = Oracle(per)
    b.set_qj(Qp)
    return a.rhs()[_interior(wall)], b.rhs()[half]


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("name", list(MIRROR_SCHEMES))
def test_slip_wall_rhs_is_the_mirror_rhs_oracle(name, axis):
    """Right-hand side next to slip walls = that of the mirrored periodic case
    (every face of an interior cell reads the same values, the sensor of the
    ghost layer included), on the initial field and on the field after two
    steps between the walls."""
    wall, Qw, per, Qp, half = mirror_pair(name, axis)
    ra, rb = _rhs_pair(wall, Qw, per, Qp, half)
    assert np.abs(ra).max() > 0
    np.testing.assert_array_equal(ra, rb)
    # evolve between the walls, mirror the result, compare again
    a = Oracle(wall)
    a.set_qj(Qw)
    a.step(2)
    Qw2 = a.get_qj()
    h = wall.order // 2
    npax = {0: 3, 1: 1, 2: 0}[axis]
    f = Qw2[h:-h, h:-h, :, h:-h]
    mir = np.flip(f, axis=npax).copy()
    mir[:, :, 1 + axis] *= -1.0
    Qp2 = np.zeros(per.state_shape())
    Qp2[h:-h, h:-h, :, h:-h] = np.concatenate([mir, f], axis=npax)
    Qp2 = cases.periodic_fill(Qp2, wall.order)
    ra, rb = _rhs_pair(wall, Qw2, per, Qp2, half)
    np.testing.assert_array_equal(ra, rb)


def stz_l1(Q, spec, t):
    """Relative L1 errors of density and pressure along the middle z line
    against Sod's exact solution (the diaphragm at Lz / 2)."""
    h = spec.order // 2
    line = Q[h:-h, spec.ny // 2, :, spec.nx // 2]
    rho = line[:, 0]
    w = line[:, 3] / rho
    p = (spec.gamma - 1.0) * (line[:, 4] - 0.5 * rho * w * w)
    n = rho.size
    zc = (np.arange(n) + 0.5) * (spec.Lz / n)
    ex = np.array([make_sod.sample((z - 0.5 * spec.Lz) / t, 1.0, 0.0, 1.0, 0.125, 0.0, 0.1) for z in zc])
    return (np.mean(np.abs(rho - ex[:, 0])) / np.mean(ex[:, 0]), np.mean(np.abs(p - ex[:, 2])) / np.mean(ex[:, 2]))


STZ_L1 = 0.02  # first-order SLAU at 256 cells; the oracle reaches 1.4 % (rho) and 1.3 % (p)


def test_stz_sod_between_walls_oracle():
    spec = cases.stz_spec(256, 4)
    o = Oracle(spec)
    o.set_state(cases.stz_state(spec))
    n = int(0.1 / spec.dt)
    o.step(n)
    Q = o.get_state()
    h = spec.order // 2
    # every line along z is the same (no transverse motion between slip walls)
    np.testing.assert_array_equal(Q[h:-h, h:-h], np.broadcast_to(Q[h:-h, h:h + 1, :, h:h + 1], Q[h:-h, h:-h].shape))
    assert np.all(Q[h:-h, h:-h, 1:3] == 0.0)
    l1_rho, l1_p = stz_l1(Q, spec, n * spec.dt)
    assert l1_rho < STZ_L1 and l1_p < STZ_L1, (l1_rho, l1_p)


@pytest.mark.parametrize("kw", [dict(scheme="Roe"), dict(visc="LES", visc_order=4)])
def test_nonperiodic_refused_where_not_built(kw):
    spec = cases.stz_spec(nz=16, nt=8, order=4)
    for k, v in kw.items():
        setattr(spec, k, v)
    with pytest.raises(RuntimeError):
        Oracle(spec)


def test_stz_case_builds_with_its_walls(tmp_path):
    """Case('3D_solver/STZ') reads the case's switches and parameters
    (SLAU variant SLAU, TVD, ORDER 2, RK3) and its boundary kinds; other 3-D
    cases keep refusing BC_*=True (their ghost fill is the periodic one)."""
    from ouxsbli_amd.case import Case
    from ouxsbli_amd.ic import initial_state
    c = Case("3D_solver/STZ", str(tmp_path / "stz"))
    c.build()
    s = c.spec
    assert s.bc == ("slip", "slip", "extrapolate")
    assert (s.order, s.scheme, s.slau_variant, s.tvd, s.rk, s.visc) == (2, "SLAU", "SLAU", "tvd", 3, "Euler")
    ref = cases.stz_spec(256, 4)
    assert (s.nx, s.ny, s.nz, s.Lx, s.Ly, s.Lz, s.dt) == (ref.nx, ref.ny, ref.nz, ref.Lx, ref.Ly, ref.Lz, ref.dt)
    np.testing.assert_array_equal(initial_state("STZ", s, c.globals), cases.stz_state(ref))
    with pytest.raises(RuntimeError):
        Case("3D_solver/KHI", str(tmp_path / "k"), bc_y=True).build()
