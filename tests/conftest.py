import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (HERE, ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libhbj.so on cuda:0)")
    config.addinivalue_line("markers", "slow: full-size (configs at 50^3) parity checks")


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib
    oracle_lib.build()
    oracle_lib.set_threads(min(16, os.cpu_count() or 1))
    return oracle_lib


@pytest.fixture(scope="session")
def hbj():
    import wgpu_solver_backend_amd as H
    H._lib.lib()  # raises if libhbj.so is missing
    return H


@pytest.fixture(scope="session")
def ctx(hbj):
    c = hbj.GpuContext.create(0)
    yield c
    c.close()
