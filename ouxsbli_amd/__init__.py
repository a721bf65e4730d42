"""ouxsbli_amd -- MI355X (gfx950) RHS evaluation + Runge-Kutta time advance of
the OUxSBLI compressible-flow solver.

The compute path is the in-tree HIP library ``libouxsbli.so`` (C ABI of
include/ouxsbli.h) loaded with ctypes; there is no CPU fallback.
"""
from .cases import CaseSpec, tgv_keep4_spec, supersonic_tgv_spec, tgv_spec, tgv_state  # noqa: F401
from .case import Case  # noqa: F401
from .patcher import patch  # noqa: F401
from .ic import UniformIC, TaylorGreenIC, RiemannIC, PeriodicBC  # noqa: F401

__version__ = "0.1.0"


def Solver(*a, **k):  # lazy: importing the package does not load the library
    from .solver import Solver as _S
    return _S(*a, **k)
