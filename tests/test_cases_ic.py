"""Initial fields of the benchmark workloads (ouxsbli_amd/cases.py): the
decaying-isotropic-turbulence field (solenoidal, spectrum peak, turbulent
Mach number, seeded) and the Kelvin-Helmholtz shear layer."""
import math

import numpy as np

from ouxsbli_amd import cases


def test_dhit_velocity_is_solenoidal_with_the_stated_spectrum():
    n = 32
    u = cases.dhit_velocity(n, seed=3)
    uh = [np.fft.fftn(c) for c in u]
    k = np.fft.fftfreq(n, 1.0 / n)
    KZ, KY, KX = np.meshgrid(k, k, k, indexing="ij")
    div = KX * uh[0] + KY * uh[1] + KZ * uh[2]
    assert np.abs(div).max() <= 1e-12 * np.abs(uh[0]).max()
    E = sum(np.abs(c) ** 2 for c in uh)
    shell = np.bincount(np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2).round().astype(int).ravel(), E.ravel())
    assert shell[0] <= 1e-20 * shell.max()          # no mean flow
    assert int(np.argmax(shell[1:12])) + 1 == 4     # E(k) ~ k^4 exp(-2 (k/4)^2) peaks at k = 4
    # the shell energies follow the spectrum's shape (ratio to k = 4 within sampling scatter)
    ratio = lambda kk: shell[kk] / shell[4]
    model = lambda kk: (kk ** 4 * math.exp(-2 * (kk / 4) ** 2)) / (4 ** 4 * math.exp(-2))
    for kk in (2, 6, 8):
        assert abs(math.log(ratio(kk) / model(kk))) < 0.5, kk


def test_dhit_state_mach_density_and_seed():
    s = cases.dhit_spec(24, "fp64", seed=11)
    Q = cases.initial_state(s)
    h = s.order // 2
    rho = Q[h:-h, h:-h, 0, h:-h]
    u = Q[h:-h, h:-h, 1:4, h:-h] / rho[:, :, None, :]
    c0 = math.sqrt(s.gamma * s.R * s.T0)
    assert abs(math.sqrt(np.mean(np.sum(u * u, axis=2))) / c0 - 0.3) < 1e-12
    assert np.all(rho == s.RHO0)
    p = (s.gamma - 1) * (Q[h:-h, h:-h, 4, h:-h] - 0.5 * rho * np.sum(u * u, axis=2))
    np.testing.assert_allclose(p, s.p0, rtol=1e-12)
    np.testing.assert_array_equal(Q, cases.initial_state(cases.dhit_spec(24, "fp64", seed=11)))
    assert not np.array_equal(Q, cases.initial_state(cases.dhit_spec(24, "fp64", seed=12)))
    np.testing.assert_array_equal(Q, cases.periodic_fill(Q.copy(), s.order))


def test_khi_shear_layer():
    s = cases.khi_spec(32, 4, "KEEP", seed=5)
    Q = cases.initial_state(s)
    h = s.order // 2
    rho = Q[h:-h, h:-h, 0, h:-h]
    u = Q[h:-h, h:-h, 1, h:-h] / rho
    v = Q[h:-h, h:-h, 2, h:-h] / rho
    U = s.V0
    y = cases.cell_centres(s.ny, s.Ly, s.order)[h:-h]
    d = s.Ly / 28
    prof = 0.5 * U * (np.tanh((y - 0.25) / d) - np.tanh((y - 0.75) / d) - 1.0)
    assert np.abs(u - prof[None, :, None]).max() <= 1e-3 * U * (1 + 1e-12)
    assert np.abs(v).max() <= 1e-3 * U * (1 + 1e-12) and np.abs(v).max() > 0.5e-3 * U
    assert abs(U - 0.5) < 1e-15  # Ma 0.5 at c0 = 1
    np.testing.assert_array_equal(Q, cases.initial_state(cases.khi_spec(32, 4, "KEEP", seed=5)))
