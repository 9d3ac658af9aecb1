import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))  # CPU oracles, the tests' checkers


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow_cpu: CPU check that takes about two minutes")


def pytest_collection_modifyitems(config, items):
    # GPU tests are never silently skipped on a machine that has a device;
    # on a CPU-only machine they are deselected by -m "not gpu".
    pass
