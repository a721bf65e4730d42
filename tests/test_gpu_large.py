"""A grid beyond 2^31 state values on one GPU.

760^3 cells with ghosts (754^3 interior, ORDER 6) hold 5 * 760^3 = 2.19e9
state values, so the state indices of the top z planes exceed 32 bits.  A
parallel shear flow u(y) (the shear-layer initial state without noise, built
on the device) stays independent of x and z: every face flux along x and z
is the same on both sides of a cell and cancels exactly.  After two RK3 steps
of the supersonic-TGV scheme (SLAU6 HR-SLAU2, hybrid TVD, central-6 viscous),
the second with segment lengths at the library's cap (the longest whose
32-bit offsets stay below 2^31), every (x, z) column of the large grid's
state, ghosts included, must equal
bit for bit the oracle's state on a grid of the same y extent and cell sizes
that is 16 cells wide in x and z.
"""
import numpy as np
import pytest

from ouxsbli_amd import _abi
from ouxsbli_amd._native import OxError
from ouxsbli_amd.batch import footprint
from ouxsbli_amd.cases import CaseSpec
from ouxsbli_amd.solver import Solver
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

N_BIG = 760           # padded extent, 754 interior cells at ORDER 6
N_SMALL = 16 + 6      # 16 interior cells in x and z for the oracle


def _spec(nx, nz):
    h = 3
    n_int = N_BIG - 2 * h
    dx = 1.0 / n_int
    # Lx = (nx - 6) / n_int keeps the cell size of the large grid bit for bit
    # when nx - 6 is a power of two (scaling by 2^k is exact)
    return CaseSpec(nx=nx, ny=N_BIG, nz=nz, order=6, scheme="SLAU", visc="NS", visc_order=6, tvd="hybrid",
                    slau_variant="HRSLAU2", rk=3, precision="fp64", threshold=0.4,
                    dt=0.3 * dx / (343.0 + 100.0), Lx=(nx - 2 * h) / n_int, Ly=1.0, Lz=(nz - 2 * h) / n_int,
                    ic="khi_tanh", V0=100.0, RHO0=1.2, p0=101325.0)


def _ic(spec):
    return _abi.OxIc(kind=_abi.OX_IC_SHEAR_LAYER, seed=0, V0=spec.V0, rho0=spec.RHO0, p0=spec.p0,
                     delta=spec.Ly / 28.0, noise=0.0)


def test_grid_over_2e31_state_values_matches_oracle_columns():
    big = _spec(N_BIG, N_BIG)
    assert 5 * big.nx * big.ny * big.nz > 2**31
    need = footprint(big)
    try:
        import psutil
        if psutil.virtual_memory().available < 40 << 30:
            pytest.skip("needs 40 GiB of free host memory")
    except ImportError:
        pass

    small = _spec(N_SMALL, N_SMALL)
    with Solver(small) as s:
        s.init_state(_ic(small))
        Q0 = s.get_qj()
        s.step(2)
        Qs = s.get_qj()
    o = Oracle(small)
    o.set_qj(Q0)
    o.step(2)
    ref = o.get_qj()
    np.testing.assert_array_equal(Qs, ref)
    # the premise: the oracle's state is the same in every (x, z) column
    col = ref[3, :, :, 3]  # (ny, 5)
    assert np.array_equal(ref, np.broadcast_to(col[None, :, :, None], ref.shape))
    assert np.abs(ref[:, :, 1, :] - Q0[:, :, 1, :]).max() > 0  # the flow moved

    print(f"large grid {big.nx}x{big.ny}x{big.nz}, {need / 2**30:.1f} GiB", flush=True)
    try:
        s = Solver(big)
    except OxError as e:
        if e.code == _abi.OX_ENOMEM:
            pytest.skip(f"needs {need / 2**30:.1f} GiB of device memory")
        raise
    with s:
        s.init_state(_ic(big))
        s.step(1)
        # the longest segments the library accepts: a z segment of 741 cells
        # reaches offsets of 2.14e9 from its column pointers
        s.set_segments(4096, 4096)
        s.step(1)
        Q = s.get_qj()
    for k0 in range(0, big.nz, 40):
        blk = Q[k0:k0 + 40]
        ok = np.array_equal(blk, np.broadcast_to(col[None, :, :, None], blk.shape))
        assert ok, f"z planes {k0}..{k0 + blk.shape[0] - 1} differ from the oracle's column"
