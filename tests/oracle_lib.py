"""ctypes loader for the ORACLE (oracle/liboxoracle.so).

Test infrastructure: the oracle is the CPU reference the HIP library is
compared against.  Builds the library on first use if it is missing (make -C oracle).
"""
import ctypes
import os
import pathlib
import subprocess

import numpy as np

from ouxsbli_amd import _abi

ROOT = pathlib.Path(__file__).resolve().parent.parent
LIB = ROOT / "oracle" / "liboxoracle.so"

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            subprocess.run(["make", "-C", str(ROOT / "oracle")], check=True, capture_output=True)
        L = ctypes.CDLL(str(LIB))
        P = ctypes.c_void_p
        D = ctypes.POINTER(ctypes.c_double)
        L.oxo_create.restype = P
        L.oxo_create.argtypes = [ctypes.POINTER(_abi.OxConfig)]
        L.oxo_destroy.argtypes = [P]
        for f in ("oxo_set_state", "oxo_get_state", "oxo_set_qj", "oxo_get_qj", "oxo_rhs"):
            getattr(L, f).argtypes = [P, D]
            getattr(L, f).restype = ctypes.c_int
        L.oxo_step.argtypes = [P, ctypes.c_int]
        L.oxo_step.restype = ctypes.c_int
        L.oxo_set_bc.argtypes = [P]
        L.oxo_set_bc.restype = ctypes.c_int
        L.oxo_get_primitives.argtypes = [P, D, D, D, D]
        L.oxo_diagnostics.argtypes = [P, ctypes.POINTER(_abi.OxDiag)]
        L.oxo_get_les.argtypes = [P, D, D]
        L.oxo_last_error.restype = ctypes.c_char_p
        L.oxo_keep_flux.argtypes = [ctypes.POINTER(_abi.OxConfig), ctypes.c_int, D, ctypes.c_int, D]
        L.oxo_slau_flux.argtypes = [ctypes.POINTER(_abi.OxConfig), ctypes.c_int, D, ctypes.c_int,
                                    ctypes.c_double, D]
        L.oxo_conv_flux.argtypes = [ctypes.POINTER(_abi.OxConfig), ctypes.c_int, D, ctypes.c_int,
                                    ctypes.c_double, D]
        L.oxo_muscl.argtypes = [ctypes.POINTER(_abi.OxConfig), ctypes.c_int, ctypes.c_int, ctypes.c_double, D,
                                D, D]
        L.oxo_wiggle.argtypes = [D]
        L.oxo_wiggle.restype = ctypes.c_double
        L.oxo2d_create.restype = P
        L.oxo2d_create.argtypes = [ctypes.POINTER(_abi.Ox2dConfig)]
        L.oxo2d_destroy.argtypes = [P]
        for f in ("oxo2d_set_state", "oxo2d_get_state"):
            getattr(L, f).argtypes = [P, D]
            getattr(L, f).restype = ctypes.c_int
        L.oxo2d_step.argtypes = [P, ctypes.c_int]
        L.oxo2d_step.restype = ctypes.c_int
        L.oxo2d_last_error.restype = ctypes.c_char_p
        L.oxo2d_face_flux.argtypes = [ctypes.POINTER(_abi.Ox2dConfig), ctypes.c_int, ctypes.c_int, ctypes.c_int, D,
                                      ctypes.c_double, D]
        _lib = L
    return _lib


def _dp(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Oracle:
    """The CPU transcription driven through its C entry points."""

    def __init__(self, spec):
        self.spec = spec
        self.cfg = spec.to_ox()
        self.h = lib().oxo_create(ctypes.byref(self.cfg))
        if not self.h:
            raise RuntimeError(lib().oxo_last_error().decode())

    def close(self):
        if self.h:
            lib().oxo_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_state(self, Q):
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        lib().oxo_set_state(self.h, _dp(Q))

    def get_state(self):
        Q = np.empty(self.spec.state_shape())
        lib().oxo_get_state(self.h, _dp(Q))
        return Q

    def set_qj(self, QJ):
        QJ = np.ascontiguousarray(QJ, dtype=np.float64)
        lib().oxo_set_qj(self.h, _dp(QJ))

    def get_qj(self):
        QJ = np.empty(self.spec.state_shape())
        lib().oxo_get_qj(self.h, _dp(QJ))
        return QJ

    def step(self, n=1):
        lib().oxo_step(self.h, int(n))

    def set_bc(self):
        """The case's ghost fill on the current state (as after a stage)."""
        lib().oxo_set_bc(self.h)

    def rhs(self):
        R = np.empty(self.spec.state_shape())
        lib().oxo_rhs(self.h, _dp(R))
        return R

    def primitives(self):
        s = self.spec
        Qp = np.empty(s.state_shape())
        T = np.empty((s.nz, s.ny, s.nx))
        mu = np.empty_like(T)
        sen = np.empty_like(T)
        lib().oxo_get_primitives(self.h, _dp(Qp), _dp(T), _dp(mu), _dp(sen))
        return Qp, T, mu, sen

    def les_fields(self):
        """(mut, qc2) of the last RHS evaluation, numpy shape (nz, ny, nx)."""
        s = self.spec
        mut = np.empty((s.nz, s.ny, s.nx))
        qc2 = np.empty_like(mut)
        lib().oxo_get_les(self.h, _dp(mut), _dp(qc2))
        return mut, qc2

    def diagnostics(self):
        d = _abi.OxDiag()
        lib().oxo_diagnostics(self.h, ctypes.byref(d))
        return d.as_dict()


class Oracle2D:
    """The 2-D oblique-shock transcription driven through its C entry points."""

    def __init__(self, spec):
        self.spec = spec
        self.cfg = spec.to_ox()
        self.h = lib().oxo2d_create(ctypes.byref(self.cfg))
        if not self.h:
            raise RuntimeError(lib().oxo2d_last_error().decode())

    def close(self):
        if self.h:
            lib().oxo2d_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_state(self, Q):
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        assert Q.shape == self.spec.state_shape()
        lib().oxo2d_set_state(self.h, _dp(Q))

    def get_state(self):
        Q = np.empty(self.spec.state_shape())
        lib().oxo2d_get_state(self.h, _dp(Q))
        return Q

    def step(self, n=1):
        lib().oxo2d_step(self.h, int(n))


def face_flux2d(spec, direction, f, n, st, fs):
    """Oracle flux of one 2-D face (see ox_oracle2d.cpp, Oracle2D::face):
    st[4][6] primitives (rho, u, v, p) at points f-2..f+3."""
    cfg = spec.to_ox()
    st = np.ascontiguousarray(st, dtype=np.float64).reshape(4, 6)
    out = np.empty(4)
    lib().oxo2d_face_flux(ctypes.byref(cfg), int(direction), int(f), int(n), _dp(st), float(fs), _dp(out))
    return out
