"""``ouxsbli``: the reference package's import name over the MI355X build.

Scripts written for the reference (``from ouxsbli import Case``,
``from ouxsbli.patcher import patch``; ouxsbli/__init__.py of the reference)
run unchanged: the names resolve to ouxsbli_amd, whose Case builds and runs a
case on the gfx950 library instead of compiling and launching the CUDA
Fortran executable.
"""
from ouxsbli_amd.case import Case  # noqa: F401
from ouxsbli_amd.patcher import patch  # noqa: F401
from ouxsbli_amd.ic import UniformIC, TaylorGreenIC, RiemannIC, PeriodicBC  # noqa: F401

__all__ = ["Case", "patch", "UniformIC", "TaylorGreenIC", "RiemannIC", "PeriodicBC"]
