"""Python handle over one device-resident solve (the C ABI of include/ouxsbli.h).

``Solver`` replaces what the reference's ``RungeKutta`` does inside its
executable (3D_solver/src/calc_time_dev.f90.fypp:30-114, :224-300): it owns
the device state, advances it by RK steps and hands back the conserved
fields.  All arithmetic runs in ``libouxsbli.so`` on the GPU.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _abi, _native
from .cases import CaseSpec


def _dp(a: np.ndarray):
    if a.dtype != np.float64 or not a.flags.c_contiguous:
        raise ValueError("expected a C-contiguous float64 array")
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Solver:
    def __init__(self, spec: CaseSpec):
        self.spec = spec
        self._lib = _native.load()
        self._cfg = spec.to_ox()
        h = ctypes.c_void_p()
        _native.check(self._lib.ox_create(ctypes.byref(self._cfg), ctypes.byref(h)))
        self._h = h

    # -- lifecycle ---------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            self._lib.ox_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state -------------------------------------------------------------
    def _check_shape(self, a: np.ndarray) -> np.ndarray:
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != self.spec.state_shape():
            raise ValueError(f"state shape {a.shape} != {self.spec.state_shape()}")
        return a

    def set_state(self, Q: np.ndarray) -> None:
        """Physical conserved state Q(nx,5,ny,nz) (numpy shape (nz, ny, 5, nx))."""
        Q = self._check_shape(Q)
        _native.check(self._lib.ox_set_state(self._h, _dp(Q)))

    def init_state(self, ic=None) -> None:
        """Builds the initial state on the device (ox_init_state) from an
        ``ox_ic``, by default the case's own (CaseSpec.device_ic)."""
        ic = ic if ic is not None else self.spec.device_ic()
        if ic is None:
            raise ValueError(f"no device initial condition for ic {self.spec.ic!r}")
        _native.check(self._lib.ox_init_state(self._h, ctypes.byref(ic)))

    def get_state(self) -> np.ndarray:
        Q = np.empty(self.spec.state_shape())
        _native.check(self._lib.ox_get_state(self._h, _dp(Q)))
        return Q

    def set_qj(self, QJ: np.ndarray) -> None:
        QJ = self._check_shape(QJ)
        _native.check(self._lib.ox_set_qj(self._h, _dp(QJ)))

    def get_qj(self) -> np.ndarray:
        QJ = np.empty(self.spec.state_shape())
        _native.check(self._lib.ox_get_qj(self._h, _dp(QJ)))
        return QJ

    # -- compute -----------------------------------------------------------
    def step(self, n: int = 1) -> float:
        """Advance ``n`` RK steps; returns the device time in milliseconds."""
        _native.check(self._lib.ox_step(self._h, int(n)))
        ms = ctypes.c_double()
        _native.check(self._lib.ox_last_step_ms(self._h, ctypes.byref(ms)))
        return ms.value

    def rhs(self) -> np.ndarray:
        R = np.empty(self.spec.state_shape())
        _native.check(self._lib.ox_rhs(self._h, _dp(R)))
        return R

    def diagnostics(self) -> dict:
        d = _abi.OxDiag()
        _native.check(self._lib.ox_diagnostics(self._h, ctypes.byref(d)))
        return d.as_dict()

    def check_state(self):
        """(ok, first_bad_linear_index, count)."""
        idx = ctypes.c_int64()
        cnt = ctypes.c_int64()
        rc = self._lib.ox_check_state(self._h, ctypes.byref(idx), ctypes.byref(cnt))
        if rc not in (_abi.OX_OK, _abi.OX_ESTATE):
            _native.check(rc)
        return rc == _abi.OX_OK, idx.value, cnt.value

    def set_profiling(self, enable: bool) -> None:
        _native.check(self._lib.ox_set_profiling(self._h, 1 if enable else 0))

    def kernel_times(self) -> dict:
        """{kernel name: (total_ms, launches)} accumulated while profiling."""
        cap = 32
        arr = (_abi.OxKernelTime * cap)()
        n = ctypes.c_int32()
        _native.check(self._lib.ox_kernel_times(self._h, arr, cap, ctypes.byref(n)))
        return {arr[a].name.decode(): (arr[a].total_ms, arr[a].launches) for a in range(min(n.value, cap))}

    def autotune(self):
        """Times candidate y/z sweep segment lengths on this grid and keeps
        the fastest; returns (lseg_y, lseg_z).  The state is unchanged."""
        ly, lz = ctypes.c_int32(), ctypes.c_int32()
        _native.check(self._lib.ox_autotune(self._h, ctypes.byref(ly), ctypes.byref(lz)))
        return ly.value, lz.value

    def set_segments(self, lseg_y: int, lseg_z: int) -> None:
        """Fixes the y/z sweep segment lengths (1..4096 cells)."""
        _native.check(self._lib.ox_set_segments(self._h, int(lseg_y), int(lseg_z)))

    def synchronize(self) -> None:
        _native.check(self._lib.ox_synchronize(self._h))
