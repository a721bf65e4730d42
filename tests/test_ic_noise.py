"""The host form of the device initial conditions (ouxsbli_amd/cases.py):
the counter-based noise generator and the shear-layer field built on it.

SplitMix64 is pinned by its published output sequence (the generator of
Steele, Lea and Flood seeded with 0); the noise map on top of it by its
exact range, its moments and the independence of the three components.
The device kernels (ouxsbli_amd/csrc/ox_ic.hip) are checked against these
host functions in tests/test_gpu_ic.py."""
import numpy as np

from ouxsbli_amd import cases


def test_splitmix64_known_sequence():
    # the first outputs of SplitMix64 seeded with 0: output k is the mix of (k + 1) * 0x9E3779B97F4A7C15
    want = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    got = cases.splitmix64(np.array([0, 0x9E3779B97F4A7C15, (2 * 0x9E3779B97F4A7C15) % (1 << 64)], dtype=np.uint64))
    assert [int(x) for x in got] == want


def test_cell_noise_range_moments_and_independence():
    shape = (20, 30, 40)
    f = [cases.cell_noise(7, shape, c) for c in range(3)]
    for a in f:
        assert a.shape == shape
        assert a.min() >= -1.0 and a.max() < 1.0
        assert abs(a.mean()) < 0.01
        assert abs(a.var() - 1.0 / 3.0) < 0.01  # uniform on [-1, 1)
        # 53-bit grid: 2 * k * 2^-53 - 1 exactly
        k = (a + 1.0) * 2.0 ** 52
        assert np.array_equal(k, np.floor(k))
    for a, b in ((0, 1), (0, 2), (1, 2)):
        assert abs(np.corrcoef(f[a].ravel(), f[b].ravel())[0, 1]) < 0.01
    # neighbouring cells are uncorrelated
    assert abs(np.corrcoef(f[0].ravel()[:-1], f[0].ravel()[1:])[0, 1]) < 0.01


def test_cell_noise_is_a_function_of_seed_field_and_cell():
    a = cases.cell_noise(3, (4, 5, 6), 1)
    b = cases.cell_noise(3, (4, 5, 6), 1)
    c = cases.cell_noise(4, (4, 5, 6), 1)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, c)
    # the same linear cell gives the same value whatever the array's shape
    d = cases.cell_noise(3, (120,), 1)
    assert np.array_equal(a.ravel(), d)


def test_shear_layer_state():
    spec = cases.khi_spec(32, 4, "SLAU", seed=11)
    Q = cases.khi_state(spec)
    h = spec.order // 2
    inner = Q[h:-h, h:-h, :, h:-h]
    rho = inner[:, :, 0, :]
    assert np.all(rho == spec.RHO0)
    u = inner[:, :, 1, :] / rho
    y = cases.cell_centres(spec.ny, spec.Ly, spec.order)[h:-h]
    d = spec.Ly / 28.0
    prof = 0.5 * spec.V0 * (np.tanh((y - 0.25) / d) - np.tanh((y - 0.75) / d) - 1.0)
    dev = u - prof[None, :, None]
    amp = 1e-3 * spec.V0
    assert np.abs(dev).max() <= amp * (1 + 1e-12)
    assert np.abs(dev).max() > 0.9 * amp
    # the layers: -U/2 outside 1/4 < y < 3/4, +U/2 inside; 0.2 = 5.6 d away from
    # both layers the profile is within 2 exp(-11.2) = 3e-5 of U/2
    far = np.abs(np.abs(y - 0.5) - 0.25) > 0.2
    inside = np.abs(y - 0.5) < 0.25
    assert np.allclose(prof[far], np.where(inside[far], 0.5, -0.5) * spec.V0, atol=3e-5 * spec.V0)
    # ghost layers are periodic images
    assert np.array_equal(Q[:h], Q[-2 * h:-h])
