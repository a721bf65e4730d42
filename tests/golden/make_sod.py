"""Writes tests/golden/sod_double.npz: the exact solution of the reference's
inviscid shock tube (ouxsbli/tests/test_st.py, case 2D_solver/ST) posed as a
periodic double Riemann problem for the 3-D periodic path.

The reference runs the Sod problem of 2D_solver/ST/mod_globals.f90 (rho0 =
1.293, p0 = rho0 R T0 with T0 = 300 K; right state rho0/8, p0/10; Lx = Re mu0 /
sqrt(rho0 p0), Re = 25000, Sutherland mu0 with C = 1.461e-6, S = 110.3) on 513
points over Lx to endT = 0.2136 Lx / sqrt(p0/rho0), dt = 0.1 Lx / (512
sqrt(p0/rho0)), and requires the density and pressure L1 errors against the
exact solution to stay below 1 %.  Here the domain is [0, 2 Lx) periodic with
the same spacing (1024 cells): the high state fills (Lx/2, 3Lx/2), so a
mirrored Sod problem sits at Lx/2 and the reference's at 3Lx/2.  Until the
two shocks meet at x = 0 (t = 0.34 Lx/c0, after endT = 0.25 Lx/c0) each half
of the domain is exactly one of them.

The exact solver is Toro's (Riemann Solvers and Numerical Methods for Fluid
Dynamics, 3rd ed., ch. 4): Newton iteration on the pressure function
f_L(p) + f_R(p) + (u_R - u_L) = 0 from the PVRS guess, then sampling of the
wave pattern at x/t.  tests/test_sod.py pins it against Toro's tabulated star
state of his Test 1.

Run: python tests/golden/make_sod.py
"""
from __future__ import annotations

import math
import pathlib

import numpy as np

GAMMA = 1.4
HERE = pathlib.Path(__file__).resolve().parent


def st_parameters():
    """2D_solver/ST/mod_globals.f90:26-44 with nx = 513 (the reference test's)."""
    R, T0, C, S = 287.03, 300.0, 1.461e-6, 110.3
    mu0 = C * T0 ** 1.5 / (T0 + S)
    Re = 25000.0
    rho0 = 1.293
    p0 = rho0 * R * T0
    rho1, p1 = 0.125 * rho0, 0.1 * p0
    Lx = Re * mu0 / math.sqrt(rho0 * p0)
    nx = 513
    dt = 0.1 * Lx / ((nx - 1) * math.sqrt(p0 / rho0))
    endT = 0.2136 * Lx / math.sqrt(p0 / rho0)
    nt = int(endT / (1.0 * dt))  # np = 1 in the reference test
    return dict(R=R, T0=T0, rho0=rho0, p0=p0, rho1=rho1, p1=p1, Lx=Lx, dt=dt, endT=endT, nt=nt, n=(nx - 1) * 2)


def _pressure_function(p, rho, pk, c):
    if p > pk:  # shock
        A = 2.0 / ((GAMMA + 1.0) * rho)
        B = (GAMMA - 1.0) / (GAMMA + 1.0) * pk
        f = (p - pk) * math.sqrt(A / (p + B))
        df = math.sqrt(A / (B + p)) * (1.0 - 0.5 * (p - pk) / (B + p))
    else:  # rarefaction
        f = 2.0 * c / (GAMMA - 1.0) * ((p / pk) ** ((GAMMA - 1.0) / (2.0 * GAMMA)) - 1.0)
        df = (1.0 / (rho * c)) * (p / pk) ** (-(GAMMA + 1.0) / (2.0 * GAMMA))
    return f, df


def star_state(rl, ul, pl, rr, ur, pr):
    """Pressure and velocity of the star region (Toro 4.2-4.3)."""
    cl, cr = math.sqrt(GAMMA * pl / rl), math.sqrt(GAMMA * pr / rr)
    p = max(1e-12, 0.5 * (pl + pr) - 0.125 * (ur - ul) * (rl + rr) * (cl + cr))  # PVRS guess
    for _ in range(100):
        fl, dfl = _pressure_function(p, rl, pl, cl)
        fr, dfr = _pressure_function(p, rr, pr, cr)
        pn = p - (fl + fr + ur - ul) / (dfl + dfr)
        pn = max(pn, 1e-12 * p)
        if abs(pn - p) <= 1e-15 * 0.5 * (pn + p):
            p = pn
            break
        p = pn
    fl, _ = _pressure_function(p, rl, pl, cl)
    fr, _ = _pressure_function(p, rr, pr, cr)
    u = 0.5 * (ul + ur) + 0.5 * (fr - fl)
    return p, u


def sample(s, rl, ul, pl, rr, ur, pr):
    """(rho, u, p) at x/t = s (Toro 4.5)."""
    g = GAMMA
    cl, cr = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    ps, us = star_state(rl, ul, pl, rr, ur, pr)
    if s <= us:  # left of the contact
        if ps > pl:
            sl = ul - cl * math.sqrt((g + 1) / (2 * g) * ps / pl + (g - 1) / (2 * g))
            if s <= sl:
                return rl, ul, pl
            r = rl * ((ps / pl + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pl + 1))
            return r, us, ps
        shl = ul - cl
        if s <= shl:
            return rl, ul, pl
        cml = cl * (ps / pl) ** ((g - 1) / (2 * g))
        stl = us - cml
        if s > stl:
            return rl * (ps / pl) ** (1 / g), us, ps
        u = 2 / (g + 1) * (cl + (g - 1) / 2 * ul + s)
        c = 2 / (g + 1) * (cl + (g - 1) / 2 * (ul - s))
        return rl * (c / cl) ** (2 / (g - 1)), u, pl * (c / cl) ** (2 * g / (g - 1))
    if ps > pr:
        sr = ur + cr * math.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g))
        if s >= sr:
            return rr, ur, pr
        r = rr * ((ps / pr + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pr + 1))
        return r, us, ps
    shr = ur + cr
    if s >= shr:
        return rr, ur, pr
    cmr = cr * (ps / pr) ** ((g - 1) / (2 * g))
    str_ = us + cmr
    if s <= str_:
        return rr * (ps / pr) ** (1 / g), us, ps
    u = 2 / (g + 1) * (-cr + (g - 1) / 2 * ur + s)
    c = 2 / (g + 1) * (cr - (g - 1) / 2 * (ur - s))
    return rr * (c / cr) ** (2 / (g - 1)), u, pr * (c / cr) ** (2 * g / (g - 1))


def double_problem(x, t, P):
    """Exact (rho, u, p) of the periodic double problem at cell centres x."""
    Lx = P["Lx"]
    hi, lo = (P["rho0"], 0.0, P["p0"]), (P["rho1"], 0.0, P["p1"])
    rho, u, p = np.empty_like(x), np.empty_like(x), np.empty_like(x)
    for n, xv in enumerate(x):
        if xv < Lx:  # low | high at Lx/2: the mirror image of high | low
            r, uu, pp = sample(-(xv - 0.5 * Lx) / t, *hi, *lo)
            uu = -uu
        else:        # high | low at 3Lx/2: the reference's problem
            r, uu, pp = sample((xv - 1.5 * Lx) / t, *hi, *lo)
        rho[n], u[n], p[n] = r, uu, pp
    return rho, u, p


def main():
    P = st_parameters()
    n = P["n"]
    dx = 2.0 * P["Lx"] / n
    x = (np.arange(n) + 0.5) * dx
    rho, u, p = double_problem(x, P["endT"], P)
    np.savez(HERE / "sod_double.npz", x=x, rho=rho, u=u, p=p,
             params=np.array([P[k] for k in ("rho0", "p0", "rho1", "p1", "Lx", "dt", "endT", "nt")]))
    print(f"wrote {HERE / 'sod_double.npz'}: {n} cells, t = {P['endT']:.6e} s, nt = {P['nt']}")


if __name__ == "__main__":
    main()
