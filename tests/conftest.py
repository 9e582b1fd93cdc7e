import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "visloc-rs_amd")
ORACLE = os.path.join(ROOT, "oracle")
for p in (PKG, ORACLE, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (runs the HIP library)")
    # Build only what is missing (the GPU box receives prebuilt libraries).
    if not os.path.exists(os.path.join(PKG, "libvba.so")):
        subprocess.run(["make", "-C", PKG, "-j8"], check=True, stdout=subprocess.DEVNULL)
    if not os.path.exists(os.path.join(ORACLE, "_build", "libba_oracle.so")):
        subprocess.run(["make", "-C", ORACLE], check=True, stdout=subprocess.DEVNULL)


@pytest.fixture(scope="session")
def device():
    import vba

    dev = vba.Device(0)
    yield dev
    dev.close()
