import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "zig-voxel-physics_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)

ORACLE_SO = os.path.join(REPO, "oracle", "_build", "libvxp_oracle.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (runs the HIP pipeline)")


def _ensure_oracle():
    src = os.path.join(REPO, "oracle", "vxp_oracle.cpp")
    if not os.path.exists(ORACLE_SO) or os.path.getmtime(ORACLE_SO) < os.path.getmtime(src):
        subprocess.run(["make", "-C", os.path.join(REPO, "oracle")], check=True,
                       capture_output=True)


@pytest.fixture(scope="session")
def oracle():
    """The CPU restatement (checker)."""
    _ensure_oracle()
    import vxp
    return vxp.oracle_lib()


@pytest.fixture(scope="session")
def product():
    """The HIP pipeline; fails loudly when libvxp.so is missing."""
    import vxp
    lib = vxp.product_lib()
    return lib
