"""Pins the oracle's limited MUSCL reconstructions (src/calc_muscl.f90:19-139)
against the published schemes, evaluated in exact rational arithmetic.

- ORDER 4 (delta4 -> MUSCL3rdMinmod): van Leer's kappa scheme with kappa =
  1/3 and the minmod-limited differences of compression factor
  b = (3 - kappa)/(1 - kappa) = 4:
      qL = q_j     + (1-kappa)/4 * mm(D-, b D+) + (1+kappa)/4 * mm(D+, b D-)
      qR = q_{j+1} - (1-kappa)/4 * mm(D++, b D+) - (1+kappa)/4 * mm(D+, b D++)
- ORDER 6 (delta6 -> MUSCL4thTVD): the fourth-order compact MUSCL TVD scheme of
  Yamamoto and Daiguji (1993): corrected differences
      D*_{i+1/2} = D_{i+1/2} - (1/6) d3_{i+1/2},
      d3_{i+1/2} = mm(D_{i-1/2}, 2D_{i+1/2}, 2D_{i+3/2})
                   - 2 mm(D_{i+1/2}, 2D_{i-1/2}, 2D_{i+3/2})
                   + mm(D_{i+3/2}, 2D_{i-1/2}, 2D_{i+1/2}),
  and with b = 4
      qL = q_j     + (1/6) [mm(D*_{j-1/2}, b D*_{j+1/2}) + 2 mm(D*_{j+1/2}, b D*_{j-1/2})]
      qR = q_{j+1} - (1/6) [mm(D*_{j+3/2}, b D*_{j+1/2}) + 2 mm(D*_{j+1/2}, b D*_{j+3/2})]
- mm is the minmod function: the argument of smallest magnitude when all
  share a sign, else 0.

The inputs are small integers chosen so that every minmod argument wins in
turn (the plain difference, each compressed one, and the zero of a sign
change), plus a brute-force sweep over every integer stencil in a small box.
A wrong compression factor (2 instead of 4), a transposed difference or a
swapped left/right weight changes the result on some of them.  The oracle
runs in FP64 on integer data, so the bar is 1 ulp-level (1e-14 relative).
"""
import ctypes
import itertools
from fractions import Fraction as Fr

import numpy as np
import pytest

from ouxsbli_amd import cases
from oracle_lib import lib


def _cfg():
    return cases.CaseSpec(nx=16, ny=16, nz=16, threshold=0.4).to_ox()


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def mm(*xs):
    if all(x > 0 for x in xs):
        return min(xs)
    if all(x < 0 for x in xs):
        return max(xs)
    return Fr(0)


def kappa_scheme(a):
    """ORDER 4: a = q_{j-1}, q_j, q_{j+1}, q_{j+2}; the face is j+1/2."""
    q = [Fr(x) for x in a]
    kappa = Fr(1, 3)
    b = (3 - kappa) / (1 - kappa)
    dm, d0, dp = q[1] - q[0], q[2] - q[1], q[3] - q[2]
    qL = q[1] + (1 - kappa) / 4 * mm(dm, b * d0) + (1 + kappa) / 4 * mm(d0, b * dm)
    qR = q[2] - (1 - kappa) / 4 * mm(dp, b * d0) - (1 + kappa) / 4 * mm(d0, b * dp)
    return qL, qR


def yamamoto_daiguji(a):
    """ORDER 6: a = q_{j-2} .. q_{j+3}; the face is j+1/2."""
    q = [Fr(x) for x in a]
    D = [q[k + 1] - q[k] for k in range(5)]  # D[0] = D_{j-3/2} ... D[4] = D_{j+5/2}

    def corrected(k):  # D*_{k} from D[k-1], D[k], D[k+1]
        d3 = mm(D[k - 1], 2 * D[k], 2 * D[k + 1]) - 2 * mm(D[k], 2 * D[k - 1], 2 * D[k + 1]) + \
            mm(D[k + 1], 2 * D[k - 1], 2 * D[k])
        return D[k] - d3 / 6

    Dm, D0, Dp = corrected(1), corrected(2), corrected(3)  # j-1/2, j+1/2, j+3/2
    b = 4
    qL = q[2] + (mm(Dm, b * D0) + 2 * mm(D0, b * Dm)) / 6
    qR = q[3] - (mm(Dp, b * D0) + 2 * mm(D0, b * Dp)) / 6
    return qL, qR


def _oracle(order, a, tvd=1, sensor=0.0):
    a = np.ascontiguousarray(a, dtype=np.float64)
    al, ar = ctypes.c_double(), ctypes.c_double()
    lib().oxo_muscl(ctypes.byref(_cfg()), order, tvd, sensor, _dp(a), ctypes.byref(al), ctypes.byref(ar))
    return al.value, ar.value


def _close(got, want):
    w = float(want)
    assert got == pytest.approx(w, rel=1e-14, abs=1e-14), (got, w)


# ORDER 4 stencils: unlimited, each compressed difference winning, extrema
KAPPA_CASES = [
    (0, 1, 2, 3),        # linear: nothing clipped
    (0, 1, 6, 7),        # D+ = 5 > 4 D- = 4: clipped to 4 D- on the left
    (0, 5, 6, 7),        # D- = 5 > 4 D+ = 4: clipped on the other side
    (0, 3, 4, 14),       # right state: D++ = 10 > 4 D+
    (0, 2, 1, 3),        # sign change: first order
    (5, 3, 3, 1),        # zero difference
    (0, -2, -9, -10),    # negative monotone, clipped
    (0, 1, 4, 5),        # D+ = 3 between 2 D- and 4 D-: factor 2 would clip, 4 does not
]


@pytest.mark.parametrize("a", KAPPA_CASES)
def test_order4_minmod_cases(a):
    al, ar = _oracle(4, a)
    qL, qR = kappa_scheme(a)
    _close(al, qL)
    _close(ar, qR)


# ORDER 6 stencils: each minmod of d3 and of the final limiter winning in turn
YD_CASES = [
    (0, 1, 2, 3, 4, 5),          # linear
    (0, 1, 3, 6, 10, 15),        # quadratic: d3 unclipped
    (0, 1, 8, 27, 64, 125),      # cubic
    (0, 1, 2, 10, 11, 12),       # jump: compressed differences win
    (0, 0, 1, 1, 2, 2),          # staircase: zero differences
    (0, 3, 1, 4, 1, 5),          # oscillation: sign changes everywhere
    (0, 1, 2, 5, 9, 10),         # 2 < D*+ / D*- < 4: factor 2 and 4 differ
    (10, 9, 7, 0, -1, -2),       # decreasing with a jump
    (0, 1, 3, 4, 6, 20),         # right state clipped by the far difference
    (0, 4, 5, 6, 7, 8),          # left far difference large
]


@pytest.mark.parametrize("a", YD_CASES)
def test_order6_limited_cases(a):
    al, ar = _oracle(6, a)
    qL, qR = yamamoto_daiguji(a)
    _close(al, qL)
    _close(ar, qR)


def test_order6_brute_force():
    """Every integer stencil with differences in -3..3 (7^5 = 16807 stencils)."""
    for D in itertools.product(range(-3, 4), repeat=5):
        a = np.concatenate([[0], np.cumsum(D)]).astype(float)
        al, ar = _oracle(6, a)
        qL, qR = yamamoto_daiguji(a)
        assert al == float(qL) or abs(al - float(qL)) <= 1e-14 * max(1.0, abs(float(qL))), (D, al, qL)
        assert ar == float(qR) or abs(ar - float(qR)) <= 1e-14 * max(1.0, abs(float(qR))), (D, ar, qR)


def test_order4_brute_force():
    for D in itertools.product(range(-4, 5), repeat=3):
        a = np.concatenate([[0], np.cumsum(D)]).astype(float)
        al, ar = _oracle(4, a)
        qL, qR = kappa_scheme(a)
        assert abs(al - float(qL)) <= 1e-14 * max(1.0, abs(float(qL))), (D, al, qL)
        assert abs(ar - float(qR)) <= 1e-14 * max(1.0, abs(float(qR))), (D, ar, qR)


def test_factor_matters_on_chosen_stencils():
    """The stencils marked above really separate compression factors 2 and 4
    (so the agreement is not an accident of data the factor cannot reach)."""
    def yd_b(a, b):
        q = [Fr(x) for x in a]
        D = [q[k + 1] - q[k] for k in range(5)]

        def corrected(k):
            d3 = mm(D[k - 1], 2 * D[k], 2 * D[k + 1]) - 2 * mm(D[k], 2 * D[k - 1], 2 * D[k + 1]) + \
                mm(D[k + 1], 2 * D[k - 1], 2 * D[k])
            return D[k] - d3 / 6
        Dm, D0 = corrected(1), corrected(2)
        return q[2] + (mm(Dm, b * D0) + 2 * mm(D0, b * Dm)) / 6
    assert yd_b((0, 1, 2, 5, 9, 10), 2) != yd_b((0, 1, 2, 5, 9, 10), 4)
    assert kappa_scheme((0, 1, 4, 5))[0] != Fr(1) + mm(Fr(1), 2 * Fr(3)) / 6 + mm(Fr(3), 2 * Fr(1)) / 3
