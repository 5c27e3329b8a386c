import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X)")
    # Build only what is missing: a checkout that already holds the built
    # libraries (e.g. a copy made on a machine with the toolchain) never
    # rebuilds.
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "liboracle.so"], check=True)
    if not os.path.exists(os.path.join(ROOT, "kate_amd", "libkate_hip.so")):
        subprocess.run(["make", "-C", os.path.join(ROOT, "kate_amd"), "-j8"], check=True)


@pytest.fixture(scope="session")
def gpu():
    import kate_amd as K
    n = K.device_count()
    if n < 1:
        pytest.fail("no GPU visible to libkate_hip.so: GPU tests must run on a GPU (there is no CPU path)")
    return n
