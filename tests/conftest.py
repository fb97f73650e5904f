import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def cpu():
    import oracle
    return oracle.Cpu()


@pytest.fixture(scope="session")
def golden():
    import numpy as np
    d = os.path.join(ROOT, "tests", "golden")
    return {n[:-4]: dict(np.load(os.path.join(d, n))) for n in os.listdir(d) if n.endswith(".npz")}


@pytest.fixture(scope="session")
def gpu_ctx():
    from firtez_dz_amd import gpu
    ctx = gpu.Context(0)
    yield ctx
    ctx.close()
