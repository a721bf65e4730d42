"""``ouxsbli.case``: the reference module name (ouxsbli/case.py) over ouxsbli_amd.case."""
from ouxsbli_amd.case import *  # noqa: F401,F403
from ouxsbli_amd.case import Case  # noqa: F401
