import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "ck-bc_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

# The shared libraries are build products (not committed); build them once
# when a fresh checkout runs the CPU suite.
if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
if not os.path.exists(os.path.join(ROOT, "ck-bc_amd", "libbcmi.so")):
    subprocess.check_call(["make", "-s", "-j4", "-C", os.path.join(ROOT, "ck-bc_amd")])


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: larger sizes")


@pytest.fixture(scope="session")
def engine():
    import bcmi
    e = bcmi.Engine()
    yield e
    e.close()
