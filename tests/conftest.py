import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("", "radio-gridders_amd", "oracle"):
    p = os.path.join(REPO, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs librg's HIP kernels)")


@pytest.fixture(scope="session")
def ctx():
    import rgamd
    c = rgamd.Context(0)
    yield c
    c.close()
