"""Python handle over one device-resident 2-D solve (ox2d_* of include/ouxsbli.h).

``Solver2D`` replaces the reference's 2-D executable around
``RungeKutta`` (2D_solver/src/calc_time_dev.f90.fypp:26-108) for the
oblique-shock case: it owns the device state, advances it by RK3 steps with
the boundary conditions applied after every stage and hands back the
conserved fields Q(nx,4,ny).  All arithmetic runs in ``libouxsbli.so``.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native
from .cases2d import CaseSpec2D


def _dp(a: np.ndarray):
    if a.dtype != np.float64 or not a.flags.c_contiguous:
        raise ValueError("expected a C-contiguous float64 array")
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Solver2D:
    def __init__(self, spec: CaseSpec2D):
        self.spec = spec
        self._lib = _native.load()
        self._cfg = spec.to_ox()
        h = ctypes.c_void_p()
        _native.check(self._lib.ox2d_create(ctypes.byref(self._cfg), ctypes.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            self._lib.ox2d_destroy(self._h)
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

    def set_state(self, Q: np.ndarray) -> None:
        """Physical conserved state, numpy shape (ny, 4, nx)."""
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        if Q.shape != self.spec.state_shape():
            raise ValueError(f"state shape {Q.shape} != {self.spec.state_shape()}")
        _native.check(self._lib.ox2d_set_state(self._h, _dp(Q)))

    def init_state(self) -> None:
        """The oblique-shock initial state of the case, built on the device
        (ox2d_init_state)."""
        _native.check(self._lib.ox2d_init_state(self._h))

    def get_state(self) -> np.ndarray:
        Q = np.empty(self.spec.state_shape(), dtype=np.float64)
        _native.check(self._lib.ox2d_get_state(self._h, _dp(Q)))
        return Q

    def step(self, nsteps: int = 1) -> None:
        _native.check(self._lib.ox2d_step(self._h, int(nsteps)))

    def last_step_ms(self) -> float:
        ms = ctypes.c_double()
        _native.check(self._lib.ox2d_last_step_ms(self._h, ctypes.byref(ms)))
        return ms.value
