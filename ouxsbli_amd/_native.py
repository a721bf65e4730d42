"""ctypes binding of the in-tree gfx950 library ``ouxsbli_amd/libouxsbli.so``.

The library is the only compute path: if it is missing, or no GPU is visible,
every entry point raises; nothing falls back to a CPU implementation.
"""
from __future__ import annotations

import ctypes
import pathlib

from . import _abi

import os

# OUXSBLI_LIB selects another build of the same library (e.g. a tuning variant).
LIB_PATH = pathlib.Path(os.environ.get("OUXSBLI_LIB") or pathlib.Path(__file__).resolve().parent / "libouxsbli.so")

_lib = None


class OxError(RuntimeError):
    """A negative status from the C ABI (``code`` holds the OX_E* value)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[ox error {code}] {msg}")
        self.code = code


def load() -> ctypes.CDLL:
    """Load the library (once) and declare every exported signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m ouxsbli_amd.build` "
            "(or __graft_entry__.build()); there is no CPU fallback")
    L = ctypes.CDLL(str(LIB_PATH))
    P = ctypes.c_void_p
    D = ctypes.POINTER(ctypes.c_double)
    I = ctypes.c_int
    L.ox_create.argtypes = [ctypes.POINTER(_abi.OxConfig), ctypes.POINTER(P)]
    L.ox_create.restype = I
    L.ox_destroy.argtypes = [P]
    L.ox_destroy.restype = None
    for f in ("ox_set_state", "ox_get_state", "ox_set_qj", "ox_get_qj", "ox_rhs"):
        getattr(L, f).argtypes = [P, D]
        getattr(L, f).restype = I
    L.ox_step.argtypes = [P, ctypes.c_int32]
    L.ox_step.restype = I
    L.ox_diagnostics.argtypes = [P, ctypes.POINTER(_abi.OxDiag)]
    L.ox_diagnostics.restype = I
    L.ox_real4_sum.argtypes = [D, ctypes.c_int64, ctypes.POINTER(ctypes.c_float)]
    L.ox_real4_sum.restype = I
    L.ox_diagnostics_noalloc.argtypes = [P, ctypes.POINTER(_abi.OxDiag)]
    L.ox_diagnostics_noalloc.restype = I
    L.ox_check_state.argtypes = [P, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
    L.ox_check_state.restype = I
    L.ox_synchronize.argtypes = [P]
    L.ox_synchronize.restype = I
    L.ox_last_step_ms.argtypes = [P, D]
    L.ox_last_step_ms.restype = I
    L.ox_set_profiling.argtypes = [P, ctypes.c_int32]
    L.ox_set_profiling.restype = I
    L.ox_kernel_times.argtypes = [P, ctypes.POINTER(_abi.OxKernelTime), ctypes.c_int32,
                                  ctypes.POINTER(ctypes.c_int32)]
    L.ox_kernel_times.restype = I
    L.ox_footprint.argtypes = [ctypes.POINTER(_abi.OxConfig), ctypes.POINTER(ctypes.c_uint64)]
    L.ox_footprint.restype = I
    L.ox_autotune.argtypes = [P, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]
    L.ox_autotune.restype = I
    L.ox_set_segments.argtypes = [P, ctypes.c_int32, ctypes.c_int32]
    L.ox_set_segments.restype = I
    L.ox_batch_create.argtypes = [ctypes.POINTER(_abi.OxBatchConfig), ctypes.POINTER(P)]
    L.ox_batch_create.restype = I
    L.ox_batch_destroy.argtypes = [P]
    L.ox_batch_destroy.restype = None
    L.ox_batch_admit.argtypes = [P, ctypes.POINTER(_abi.OxConfig), D, ctypes.c_int32, ctypes.POINTER(ctypes.c_int64)]
    L.ox_batch_admit.restype = I
    L.ox_batch_tune.argtypes = [P, ctypes.POINTER(_abi.OxConfig), D]
    L.ox_batch_tune.restype = I
    L.ox_batch_step.argtypes = [P, ctypes.POINTER(ctypes.c_int32)]
    L.ox_batch_step.restype = I
    L.ox_batch_run.argtypes = [P, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]
    L.ox_batch_run.restype = I
    L.ox_batch_done.argtypes = [P, ctypes.POINTER(ctypes.c_int64), ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
    L.ox_batch_done.restype = I
    L.ox_batch_retire.argtypes = [P, ctypes.c_int64, D]
    L.ox_batch_retire.restype = I
    L.ox_batch_query.argtypes = [P, ctypes.POINTER(_abi.OxBatchStats)]
    L.ox_batch_query.restype = I
    L.ox2d_create.argtypes = [ctypes.POINTER(_abi.Ox2dConfig), ctypes.POINTER(P)]
    L.ox2d_create.restype = I
    L.ox2d_destroy.argtypes = [P]
    L.ox2d_destroy.restype = None
    for f in ("ox2d_set_state", "ox2d_get_state", "ox2d_last_step_ms"):
        getattr(L, f).argtypes = [P, D]
        getattr(L, f).restype = I
    L.ox2d_step.argtypes = [P, ctypes.c_int32]
    L.ox2d_step.restype = I
    L.ox2d_footprint.argtypes = [ctypes.POINTER(_abi.Ox2dConfig), ctypes.POINTER(ctypes.c_uint64)]
    L.ox2d_footprint.restype = I
    L.ox_batch_admit2d.argtypes = [P, ctypes.POINTER(_abi.Ox2dConfig), D, ctypes.c_int32,
                                   ctypes.POINTER(ctypes.c_int64)]
    L.ox_batch_admit2d.restype = I
    L.ox_init_state.argtypes = [P, ctypes.POINTER(_abi.OxIc)]
    L.ox_init_state.restype = I
    L.ox_batch_admit_ic.argtypes = [P, ctypes.POINTER(_abi.OxConfig), ctypes.POINTER(_abi.OxIc), ctypes.c_int32,
                                    ctypes.POINTER(ctypes.c_int64)]
    L.ox_batch_admit_ic.restype = I
    L.ox2d_init_state.argtypes = [P]
    L.ox2d_init_state.restype = I
    L.ox_host_alloc.argtypes = [ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p)]
    L.ox_host_alloc.restype = I
    L.ox_host_free.argtypes = [P]
    L.ox_host_free.restype = I
    L.ox_last_error.argtypes = []
    L.ox_last_error.restype = ctypes.c_char_p
    L.ox_version.argtypes = []
    L.ox_version.restype = ctypes.c_char_p
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != _abi.OX_OK:
        raise OxError(rc, load().ox_last_error().decode(errors="replace"))
