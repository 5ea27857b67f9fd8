import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libhdrgpu.so")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle_py
    return oracle_py


@pytest.fixture(scope="session")
def gpu_ctx():
    from hdrestir_amd import gpu
    ctx = gpu.Context(0)
    yield ctx
    ctx.close()
