import os
import sys
import pathlib

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

# Under pytest-xdist every worker loads its own copy of the OpenMP oracle; left
# at one thread per core each, the workers oversubscribe the host and the
# small-grid oracle runs spend their time in OpenMP barriers (minutes instead
# of a second).  Give each worker its share of the cores before the oracle
# library loads.
if "PYTEST_XDIST_WORKER_COUNT" in os.environ and "OMP_NUM_THREADS" not in os.environ:
    _workers = max(1, int(os.environ["PYTEST_XDIST_WORKER_COUNT"]))
    os.environ["OMP_NUM_THREADS"] = str(max(1, (os.cpu_count() or 1) // _workers))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libouxsbli.so")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def gpu_available() -> bool:
    try:
        import torch  # noqa: F401  (device enumeration only)
        return torch.cuda.is_available()
    except Exception:
        return False
