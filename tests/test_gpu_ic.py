"""Initial states built on the device (ox_init_state, ox2d_init_state,
ox_batch_admit_ic, ox_batch_admit2d without a state).

The bar: each equals its host form in ouxsbli_amd/cases.py and cases2d.py
(bit for bit where no transcendental function is involved, to 1e-14 of the
field's scale where the device's sin, cos or tanh round differently from the
host's); the handle after ox_init_state behaves exactly as after uploading
the same state (RK4 accumulator zero, primitives rebuilt); and a case
admitted to a batch from its parameters ends bit-identical to the same case
advanced alone from the same device-built state."""
import dataclasses
import math

import numpy as np
import pytest

from ouxsbli_amd import cases, cases2d as c2, ic
from ouxsbli_amd.cases import CaseSpec
from ouxsbli_amd.batch import Batch
from ouxsbli_amd.solver import Solver
from ouxsbli_amd.solver2d import Solver2D

pytestmark = pytest.mark.gpu

GIB = 1 << 30


def _device_state(spec):
    with Solver(spec) as s:
        s.init_state()
        return s.get_state(), s.get_qj()


def _uploaded(spec, Q):
    """Q through the upload path and back (Q / J on the device, then J * QJ),
    which is what the device-built state is compared with."""
    with Solver(spec) as s:
        s.set_state(Q)
        return s.get_state()


def _close(a, b, tol):
    for l in range(a.shape[2]):
        scale = np.abs(b[:, :, l, :]).max()
        if scale == 0.0:
            assert np.array_equal(a[:, :, l, :], b[:, :, l, :])
        else:
            assert np.abs(a[:, :, l, :] - b[:, :, l, :]).max() <= tol * scale, f"variable {l}"


@pytest.mark.parametrize("spec", [cases.supersonic_tgv_spec(40), cases.tgv_keep4_spec(24),
                                  cases.tgv_spec(30, 4, Ma=0.3, scheme="KEEP", visc="Euler", rk=3)],
                         ids=["supersonic-tgv-40", "tgv-keep4-24", "euler-30"])
def test_taylor_green_matches_host(spec):
    print(f"device TGV {spec.interior} ORDER {spec.order}", flush=True)
    got, _ = _device_state(spec)
    _close(got, _uploaded(spec, cases.tgv_state(spec)), 1e-14)


@pytest.mark.parametrize("n,order,seed", [(32, 4, 3), (48, 6, 20261207)])
def test_shear_layer_matches_host(n, order, seed):
    print(f"device shear layer {n} ORDER {order} seed {seed}", flush=True)
    spec = cases.khi_spec(n, order, "SLAU", seed=seed)
    got, _ = _device_state(spec)
    ref = _uploaded(spec, cases.khi_state(spec))
    # rho, rho v, rho w hold no transcendental function: bit for bit
    for l in (0, 2, 3):
        assert np.array_equal(got[:, :, l, :], ref[:, :, l, :]), f"variable {l}"
    _close(got, ref, 1e-14)


def _ref_case_spec(n, order, L):
    (nx, ny, nz), (Lx, Ly, Lz) = n, L
    return CaseSpec(nx=nx + order, ny=ny + order, nz=nz + order, order=order, scheme="KEEP", visc="Euler",
                    tvd="none", rk=3, Lx=Lx, Ly=Ly, Lz=Lz)


def _device_case_state(spec, case, g):
    with Solver(spec) as s:
        s.init_state(ic.device_ic(case, spec, g))
        return s.get_state()


@pytest.mark.parametrize("order", [4, 6])
def test_etgv_matches_host(order):
    """ETGV's set_init on the device (OX_IC_ETGV) against its host form."""
    spec = _ref_case_spec((24, 20, 28), order, (2 * math.pi,) * 3)
    g = {"rho0": 1.0, "m0": 0.1}
    got = _device_case_state(spec, "ETGV", g)
    ref = _uploaded(spec, ic.etgv_state(spec, g))
    for l in (0, 3):  # rho, rho w: no transcendental function
        assert np.array_equal(got[:, :, l, :], ref[:, :, l, :]), f"variable {l}"
    _close(got, ref, 1e-14)


@pytest.mark.parametrize("order", [4, 6])
def test_khi_matches_host(order):
    """KHI's set_init on the device (OX_IC_KHI) against its host form, with
    Ly = 1.6 Lx: the stream test reads Lx, as the reference writes it."""
    spec = _ref_case_spec((32, 52, 24), order, (1.0, 1.6, 0.8))
    g = {"u1": 0.5, "rho1": 1.0, "u2": -0.5, "rho2": 2.0, "p": 2.5, "amp": 0.01}
    got = _device_case_state(spec, "KHI", g)
    ref = _uploaded(spec, ic.khi_state(spec, g))
    for l in (0, 1):  # rho, rho u: the streams' constants
        assert np.array_equal(got[:, :, l, :], ref[:, :, l, :]), f"variable {l}"
    _close(got, ref, 1e-14)


def test_oblique_shock_matches_host_bitwise():
    spec = c2.os_spec(97, 49, scheme="Roe", cfl=0.3)
    with Solver2D(spec) as s:
        s.init_state()
        got = s.get_state()
    ref = c2.os_state(spec)
    with Solver2D(spec) as s:
        s.set_state(ref)
        up = s.get_state()
    assert np.array_equal(got, up)


@pytest.mark.parametrize("spec", [cases.tgv_keep4_spec(16), cases.khi_spec(24, 6, "KEEP", seed=5)],
                         ids=["rk4-keep4-ns", "rk3-keep6-euler"])
def test_init_state_leaves_the_handle_as_an_upload_does(spec):
    """Run A from ox_init_state, run B from ox_set_qj of A's initial QJ, both
    after a previous run on the handle (so stale primitives and a nonzero RK4
    accumulator would show): identical after 4 steps."""
    with Solver(spec) as a, Solver(spec) as b:
        a.init_state()
        QJ0 = a.get_qj()
        for s in (a, b):  # dirty both handles first
            s.set_qj(QJ0 * 1.001)
            s.step(2)
        a.init_state()
        b.set_qj(QJ0)
        a.step(4)
        b.step(4)
        assert np.array_equal(a.get_qj(), b.get_qj())


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_admits_from_parameters(use_graph):
    """Cases admitted without a state (3-D TGV and shear layer, a 2-D oblique
    shock) end bit-identical to the same cases advanced alone from their
    device-built states."""
    mix = [(cases.khi_spec(32, 4, "SLAU", "fp32_bulk", seed=8), 5),
           (cases.tgv_spec(24, 6, Ma=0.5, scheme="SLAU", visc="NS", rk=3, tvd="hybrid"), 4),
           (c2.os_spec(150, 70, scheme="Roe", precision="fp32_bulk", cfl=0.3), 6)]
    ref = []
    for spec, n in mix:
        if getattr(spec, "dim", 3) == 2:
            with Solver2D(spec) as s:
                s.init_state()
                s.step(n)
                ref.append(s.get_state())
        else:
            with Solver(spec) as s:
                s.init_state()
                s.step(n)
                ref.append(s.get_state())
    with Batch(hbm_budget=4 * GIB, max_cases=8, use_graph=use_graph, autotune=False) as b:
        ids = {b.admit(spec, None, n): k for k, (spec, n) in enumerate(mix)}
        out = {}
        while len(out) < len(mix):
            b.step()
            for cid in b.done():
                out[ids[cid]] = b.retire(cid)
    for k in range(len(mix)):
        assert np.array_equal(out[k], ref[k]), f"case {k}"


def test_device_ic_is_refused_without_parameters():
    spec = cases.dhit_spec(16, "fp64")
    spec.ic = "host-only"  # an initial condition the library has no device form of
    assert spec.device_ic() is None
    with Solver(spec) as s:
        with pytest.raises(ValueError):
            s.init_state()


@pytest.mark.parametrize("n", [24, 64, 96])
def test_isotropic_spectrum_matches_host_form(n):
    """The LES case's isotropic field synthesised on the device from its mode box
    (OX_IC_SPECTRAL) against the host form (inverse real FFT of the whole half
    spectrum, cases.dhit_state): within 1e-12 of each field's scale.  At
    n = 64 and 96 the box (|k| <= 28) leaves out modes the FFT keeps; their
    amplitudes are below 1e-20 of the largest.  The run is then the same
    solve: 2 steps from either state agree to 1e-11."""
    spec = cases.dhit_spec(n, "fp64", seed=11)
    Qd, QJd = _device_state(spec)
    Qh = _uploaded(spec, cases.dhit_state(spec))
    _close(Qd, Qh, 1e-12)
    with Solver(spec) as s:
        s.set_qj(QJd)
        s.step(2)
        a = s.get_state()
    with Solver(spec) as s:
        s.set_state(cases.dhit_state(spec))
        s.step(2)
        b = s.get_state()
    _close(a, b, 1e-11)


def test_isotropic_spectrum_properties():
    """The device field is solenoidal to the accuracy of the spectral
    synthesis (the discrete spectral divergence of the box vanishes) and has
    the requested rms: sqrt(<|u|^2>) = V0."""
    spec = cases.dhit_spec(48, "fp64", seed=3)
    Qd, _ = _device_state(spec)
    h = spec.order // 2
    q = Qd[h:-h, h:-h, :, h:-h]
    u = q[:, :, 1:4, :] / q[:, :, 0:1, :]
    rms = np.sqrt(np.mean(np.sum(u * u, axis=2)))
    assert rms == pytest.approx(spec.V0, rel=1e-13)
    # spectral divergence: i k . u_hat over the FFT of the interior field
    n = spec.nx - spec.order
    U = [np.fft.fftn(u[:, :, c, :]) for c in range(3)]  # (z, y, x)
    k = np.fft.fftfreq(n, d=1.0 / n)
    KZ, KY, KX = np.meshgrid(k, k, k, indexing="ij")
    div = KX * U[0] + KY * U[1] + KZ * U[2]
    assert np.abs(div).max() <= 1e-10 * np.abs(U[0]).max() * n
