import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "rod-dynamics-3d_amd")
ORACLE = os.path.join(ROOT, "oracle")
for p in (PKG, ORACLE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs librdk kernels)")


@pytest.fixture(scope="session")
def gpu_ctx_factory():
    from rod_dynamics_py import Context

    made = []

    def make():
        c = Context(0)
        made.append(c)
        return c

    yield make
    for c in made:
        c.close()
