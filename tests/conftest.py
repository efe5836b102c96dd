import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "gvdb-voxels_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); parity tests of the HIP path")


def gpu_available():
    try:
        import gvdb_amd
        gvdb_amd.load()
        v = gvdb_amd.VolumeGVDB(0)
        v.close()
        return True
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    import gvdb_amd
    # no fallback: a GPU test must fail loudly when the HIP library or device is missing
    gvdb_amd.load()
    return gvdb_amd
