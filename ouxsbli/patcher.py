"""``ouxsbli.patcher``: the reference module name (ouxsbli/patcher.py) over ouxsbli_amd.patcher."""
from ouxsbli_amd.patcher import *  # noqa: F401,F403
from ouxsbli_amd.patcher import patch  # noqa: F401
