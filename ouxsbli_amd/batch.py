"""Python handle over the batch engine of include/ouxsbli.h (ox_batch_*).

Many cases are resident on the GPU at once; every ``step()`` advances all of
them by one RK step, with one persistent launch per sweep and RK stage for
each group of cases that share a kernel specialisation, replayed from a
captured HIP graph while the set of cases is unchanged.  Cases join with
``admit`` (refused with ``Busy`` while they do not fit in the free slots or
HBM budget) and leave when their step count is reached (``done`` /
``retire``).  The reference has no counterpart: it runs one case per
``mpirun`` (ouxsbli/case.py:180-201).
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

from . import _abi, _native
from .cases import CaseSpec


class Busy(Exception):
    """Admission refused for now (no free slot or HBM budget)."""


def footprint(spec) -> int:
    """Device bytes a case of this specification (3-D CaseSpec or 2-D
    CaseSpec2D) holds while resident."""
    b = ctypes.c_uint64()
    cfg = spec.to_ox()
    L = _native.load()
    fn = L.ox2d_footprint if getattr(spec, "dim", 3) == 2 else L.ox_footprint
    _native.check(fn(ctypes.byref(cfg), ctypes.byref(b)))
    return int(b.value)


class PinnedPool:
    """Page-locked host buffers (ox_host_alloc) for results, reused by size.

    ``array(shape)`` returns a float64 numpy array on a pinned buffer; when
    the array (and every view of it) is gone, the buffer goes back to the
    pool for the next result of the same size, up to ``keep_bytes`` retained
    in all; beyond that it is freed."""

    def __init__(self, keep_bytes: int = 8 << 30):
        self._lib = _native.load()
        self._free = {}
        self._kept = 0
        self.keep_bytes = keep_bytes

    def array(self, shape) -> np.ndarray:
        n = int(np.prod(shape))
        nbytes = 8 * n
        stack = self._free.get(nbytes)
        if stack:
            ptr = stack.pop()
            self._kept -= nbytes
        else:
            p = ctypes.c_void_p()
            _native.check(self._lib.ox_host_alloc(nbytes, ctypes.byref(p)))
            ptr = p.value
        buf = (ctypes.c_double * n).from_address(ptr)
        buf._release = _Release(self, ptr, nbytes)  # returns the buffer when the last view is gone
        return np.frombuffer(buf, dtype=np.float64).reshape(shape)

    def _give_back(self, ptr: int, nbytes: int) -> None:
        if self._kept + nbytes <= self.keep_bytes:
            self._free.setdefault(nbytes, []).append(ptr)
            self._kept += nbytes
        else:
            self._lib.ox_host_free(ctypes.c_void_p(ptr))


class _Release:
    __slots__ = ("pool", "ptr", "nbytes")

    def __init__(self, pool, ptr, nbytes):
        self.pool, self.ptr, self.nbytes = pool, ptr, nbytes

    def __del__(self):
        try:
            self.pool._give_back(self.ptr, self.nbytes)
        except Exception:  # interpreter shutdown: the process frees its pinned memory
            pass


_POOL: Optional[PinnedPool] = None


def pinned_pool() -> PinnedPool:
    """The process's pool of pinned result buffers."""
    global _POOL
    if _POOL is None:
        _POOL = PinnedPool()
    return _POOL


class Batch:
    def __init__(self, hbm_budget: int, max_cases: int = 64, use_graph: bool = True, autotune: bool = True,
                 device: int = 0, pinned_results: bool = True):
        """``pinned_results``: retire() returns arrays on page-locked host
        buffers from the process's pool (a faster device-to-host copy)."""
        self._lib = _native.load()
        c = _abi.OxBatchConfig()
        c.abi_version = _abi.OX_ABI_VERSION
        c.device = device
        c.max_cases = max_cases
        c.use_graph = int(use_graph)
        c.autotune = int(autotune)
        c.hbm_budget = int(hbm_budget)
        h = ctypes.c_void_p()
        _native.check(self._lib.ox_batch_create(ctypes.byref(c), ctypes.byref(h)))
        self._h = h
        self._shapes = {}
        self._pinned = pinned_results

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            self._lib.ox_batch_destroy(self._h)
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

    def admit(self, spec, Q: Optional[np.ndarray], nsteps: int) -> int:
        """Admits a case for nsteps RK steps and returns its id: a 3-D
        CaseSpec with physical state Q of numpy shape (nz, ny, 5, nx), or a
        2-D CaseSpec2D with Q of shape (ny, 4, nx).  Q None builds the case's
        initial state on the device instead (CaseSpec.device_ic, or the
        oblique-shock state of a 2-D case), so nothing is uploaded.  Raises
        Busy if it does not fit now."""
        cfg = spec.to_ox()
        cid = ctypes.c_int64()
        two_d = getattr(spec, "dim", 3) == 2
        if Q is None:
            if two_d:
                rc = self._lib.ox_batch_admit2d(self._h, ctypes.byref(cfg), None, int(nsteps), ctypes.byref(cid))
            else:
                ic = spec.device_ic()
                if ic is None:
                    raise ValueError(f"no device initial condition for ic {spec.ic!r}")
                rc = self._lib.ox_batch_admit_ic(self._h, ctypes.byref(cfg), ctypes.byref(ic), int(nsteps),
                                                 ctypes.byref(cid))
        else:
            Q = np.ascontiguousarray(Q, dtype=np.float64)
            if Q.shape != spec.state_shape():
                raise ValueError(f"state shape {Q.shape} != {spec.state_shape()}")
            fn = self._lib.ox_batch_admit2d if two_d else self._lib.ox_batch_admit
            rc = fn(self._h, ctypes.byref(cfg), Q.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), int(nsteps),
                    ctypes.byref(cid))
        if rc == _abi.OX_EBUSY:
            raise Busy(self._lib.ox_last_error().decode())
        _native.check(rc)
        self._shapes[cid.value] = spec.state_shape()
        return cid.value

    def tune(self, spec: CaseSpec, Q: np.ndarray) -> None:
        """Tunes (and caches) the segment lengths of cases like ``spec``
        without admitting anything."""
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        cfg = spec.to_ox()
        _native.check(self._lib.ox_batch_tune(self._h, ctypes.byref(cfg),
                                              Q.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))

    def step(self) -> int:
        """One RK step of every active case; returns how many are still stepping."""
        n = ctypes.c_int32()
        _native.check(self._lib.ox_batch_step(self._h, ctypes.byref(n)))
        return n.value

    def run(self, max_steps: int):
        """Up to max_steps steps back to back with one synchronization; stops
        after the first step in which a case finishes.  Returns (steps run,
        cases still stepping)."""
        k = ctypes.c_int32()
        n = ctypes.c_int32()
        _native.check(self._lib.ox_batch_run(self._h, int(max_steps), ctypes.byref(k), ctypes.byref(n)))
        return k.value, n.value

    def done(self) -> list:
        cap = 1024
        ids = (ctypes.c_int64 * cap)()
        n = ctypes.c_int32()
        _native.check(self._lib.ox_batch_done(self._h, ids, cap, ctypes.byref(n)))
        return [ids[k] for k in range(n.value)]

    def retire(self, cid: int) -> np.ndarray:
        """Final physical state of a finished case (ghost layers filled); frees it."""
        shape = self._shapes[cid]
        Q = pinned_pool().array(shape) if self._pinned else np.empty(shape, dtype=np.float64)
        _native.check(self._lib.ox_batch_retire(self._h, int(cid), Q.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        del self._shapes[cid]
        return Q

    def stats(self) -> dict:
        st = _abi.OxBatchStats()
        _native.check(self._lib.ox_batch_query(self._h, ctypes.byref(st)))
        return st.as_dict()
