"""Pins the oracle's selective mixed-scale SGS model (calc_les.f90:12-125)
against closed forms.

On a velocity field that is a cubic polynomial, every discrete operator the
model applies is an exact polynomial identity:

- the central difference 0.5 (f(x+h) - f(x-h)) / h equals f' + h^2/6 f''';
- the second difference f(x+h) - 2 f(x) + f(x-h) equals h^2 f'', so the test
  filter (calc_les.f90:74-81) is f + h^2/4 (f_xx + f_yy + f_zz) with each
  direction's own spacing, and the stride filter applies it at the 27 points
  of the centre's neighbourhood.

With those, sympy gives the strain and vorticity of the field and of its
filtered version at a cell centre, and from them, by the published
definitions rather than the reference's expression tree:

- qc2 = |u - u_test|^2 / 2 at the centre;
- the angle theta between the vorticity and the filtered vorticity (arccos of
  the normalised dot product), the selector f = min(1, (tan^2(theta/2) /
  tan^2(10 deg))^2) (Sagaut's selective function on a 10 degree threshold);
- S2 = 2 S_ij S_ij from the discrete gradients;
- the filter width of Scotti, Meneveau and Lilly (1993),
  (dx dy dz)^(1/3) cosh(sqrt(4/27 (ln^2 a1 - ln a1 ln a2 + ln^2 a2))) with
  a1, a2 the two smaller spacings over the largest;
- nu_t = 0.06 f S2^(alpha/2) qc2^((1-alpha)/2) Delta^(1+alpha), alpha = 1/2;
  mut = rho nu_t.

A wrong model constant (Cm, alpha, the 10 degree angle, the 1/4 filter
weight, the 1/2 in qc2, the factor 2 of the diagonal strain) moves mut by
percent or more; the bar here is 1e-9 relative.  No GPU.
"""
import math

import numpy as np
import pytest
import sympy as sp

from ouxsbli_amd.cases import CaseSpec, cell_centres, periodic_fill
from oracle_lib import Oracle

X, Y, Z = sp.symbols("x y z")
N = (14, 14, 14)
L = (0.8, 1.0, 1.3)          # z has the largest spacing (the reference's a1, a2 are x and y over it)
X0, Y0, Z0 = 0.4, 0.5, 0.62


def _fields():
    x, y, z = X - X0, Y - Y0, Z - Z0
    u = 20.0 * y ** 3 + 0.6 * z + 0.8 * x * z ** 2 - 0.05 * x
    v = 0.4 * x + 1.5 * x ** 2 * y + 0.03 * y
    w = 0.2 * y + 2.0 * x ** 3 + 0.7 * y ** 2 * z + 0.02 * z
    return [u, v, w]


def _rho():
    return 1.2 + 0.3 * (X - X0) - 0.2 * (Y - Y0)


def _spec():
    order = 6
    s = CaseSpec(nx=N[0] + order, ny=N[1] + order, nz=N[2] + order, order=order, scheme="KEEP", visc="LES",
                 visc_order=6, tvd="none", rk=3, dt=1e-4, Lx=L[0], Ly=L[1], Lz=L[2], gamma=1.4, R=287.03)
    return s


def _state(spec):
    xc = cell_centres(spec.nx, spec.Lx, spec.order)
    yc = cell_centres(spec.ny, spec.Ly, spec.order)
    zc = cell_centres(spec.nz, spec.Lz, spec.order)
    Zg, Yg, Xg = np.meshgrid(zc, yc, xc, indexing="ij")
    fu = [sp.lambdify((X, Y, Z), f, "numpy") for f in _fields()]
    rho = sp.lambdify((X, Y, Z), _rho(), "numpy")(Xg, Yg, Zg)
    u, v, w = (f(Xg, Yg, Zg) * np.ones_like(Xg) for f in fu)
    p = 2.0e5
    Q = np.empty(spec.state_shape())
    Q[:, :, 0, :] = rho
    Q[:, :, 1, :] = rho * u
    Q[:, :, 2, :] = rho * v
    Q[:, :, 3, :] = rho * w
    Q[:, :, 4, :] = p / (spec.gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
    return periodic_fill(Q, spec.order), (xc, yc, zc)


def _closed_form(p, h):
    hx, hy, hz = h
    U = _fields()
    lap = lambda f: sp.Rational(1, 4) * (hx ** 2 * sp.diff(f, X, 2) + hy ** 2 * sp.diff(f, Y, 2) +
                                         hz ** 2 * sp.diff(f, Z, 2))
    Ut = [f + lap(f) for f in U]
    D = lambda f, var, hh: sp.diff(f, var) + hh ** 2 / 6 * sp.diff(f, var, 3)
    at = {X: p[0], Y: p[1], Z: p[2]}
    grad = lambda F: np.array([[float(D(f, var, hh).subs(at)) for var, hh in ((X, hx), (Y, hy), (Z, hz))]
                               for f in F])  # grad[c][d] = d u_c / d x_d
    g, gt = grad(U), grad(Ut)
    curl = lambda a: np.array([a[2][1] - a[1][2], a[0][2] - a[2][0], a[1][0] - a[0][1]])
    om, omt = curl(g), curl(gt)
    cos_t = float(om @ omt) / (np.linalg.norm(om) * np.linalg.norm(omt))
    theta = math.acos(max(-1.0, min(1.0, cos_t)))
    ftheta = min(1.0, (math.tan(theta / 2) ** 2 / math.tan(math.radians(10.0)) ** 2) ** 2)
    S = 0.5 * (g + g.T)
    S2 = 2.0 * float(np.sum(S * S))
    du = np.array([float((f - ft).subs(at)) for f, ft in zip(U, Ut)])
    qc2 = 0.5 * float(du @ du)
    hs = sorted(h)
    a1, a2 = hs[0] / hs[2], hs[1] / hs[2]
    la1, la2 = math.log(a1), math.log(a2)
    delta = (hx * hy * hz) ** (1.0 / 3.0) * math.cosh(math.sqrt(4.0 / 27.0 * (la1 ** 2 - la1 * la2 + la2 ** 2)))
    alpha = 0.5
    nut = 0.06 * ftheta * S2 ** (alpha / 2) * qc2 ** ((1 - alpha) / 2) * delta ** (1 + alpha)
    rho = float(_rho().subs(at))
    return rho * nut, qc2, math.degrees(theta), ftheta


def test_sms_model_on_cubic_field():
    spec = _spec()
    Q, (xc, yc, zc) = _state(spec)
    o = Oracle(spec)
    o.set_state(Q)
    o.rhs()
    mut, qc2 = o.les_fields()
    h = tuple(Li / n for Li, n in zip(L, N))
    hh = spec.order // 2
    cells = [(i, j, k) for i in range(hh + 2, hh + N[0] - 2, 3) for j in range(hh + 2, hh + N[1] - 2, 3)
             for k in range(hh + 2, hh + N[2] - 2, 4)]
    partial = 0
    for (i, j, k) in cells:
        mut_ref, qc2_ref, theta, ftheta = _closed_form((xc[i], yc[j], zc[k]), h)
        assert qc2_ref > 0
        assert qc2[k, j, i] == pytest.approx(qc2_ref, rel=1e-10), (i, j, k)
        assert mut[k, j, i] == pytest.approx(mut_ref, rel=1e-9, abs=1e-300), (i, j, k, theta, ftheta)
        partial += 0.0 < ftheta < 1.0
    # the selector is in its sensitive range (neither 0 nor clamped) at most cells
    assert partial >= len(cells) // 2


def test_sms_selector_clamps_and_vanishes():
    """Where the filtered vorticity is parallel to the vorticity the selector
    gives nu_t = 0 up to the rounding of the recovered velocities (a linear
    shear, which the test filter leaves unchanged), far below the values of
    the cubic field, where it is nonzero everywhere."""
    spec = _spec()
    Q, _ = _state(spec)
    # the linear shear u = a y alone: test-filtered vorticity equals the vorticity
    h = spec.order // 2
    yc = cell_centres(spec.ny, spec.Ly, spec.order)
    Q2 = Q.copy()
    rho = Q2[:, :, 0, :]
    u = 0.7 * yc[None, :, None] * np.ones_like(rho)
    Q2[:, :, 1, :] = rho * u
    Q2[:, :, 2, :] = 0.0
    Q2[:, :, 3, :] = 0.0
    Q2[:, :, 4, :] = 2.0e5 / 0.4 + 0.5 * rho * u * u
    o = Oracle(spec)
    o.set_state(periodic_fill(Q2, spec.order))
    o.rhs()
    mut, _ = o.les_fields()
    c = (slice(h + 2, h + N[2] - 2), slice(h + 2, h + N[1] - 2), slice(h + 2, h + N[0] - 2))
    shear = mut[c].copy()
    o.set_state(Q)
    o.rhs()
    mut, _ = o.les_fields()
    assert np.all(mut[c] > 0.0)
    assert np.abs(shear).max() < 1e-20 * mut[c].min()
