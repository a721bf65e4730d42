"""``ouxsbli.ic``: the reference module name (ouxsbli/ic.py) over ouxsbli_amd.ic."""
from ouxsbli_amd.ic import *  # noqa: F401,F403
from ouxsbli_amd.ic import UniformIC, TaylorGreenIC, RiemannIC, PeriodicBC  # noqa: F401
