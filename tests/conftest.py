import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


@pytest.fixture(scope="session")
def oracle():
    """The CPU oracle (oracle/liboracle.so), the checker for the HIP path."""
    from dem_mbd_cuda_amd._abi import Library

    path = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(path):
        pytest.fail(f"{path} is not built; run __graft_entry__.build()")
    return Library(path, "ora_")


@pytest.fixture(scope="session")
def hip():
    """libdemhip.so on the GPU; fails loudly when the extension is missing."""
    import dem_mbd_cuda_amd as dm

    lib = dm.library()
    if not dm.backend_available():
        pytest.fail("no HIP device is visible to libdemhip.so")
    return lib


def golden(name):
    from dem_mbd_cuda_amd.scene import Scene

    return Scene.load(os.path.join(GOLDEN, name))
