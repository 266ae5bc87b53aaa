import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    oracle_so = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(oracle_so):
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import build as oracle_build  # oracle/build.py
        oracle_build.build_oracle()


@pytest.fixture(scope="session")
def gpu_ready():
    import torch  # noqa: F401  (torch only for the device check)
    from nmhd_amd import capi
    capi.load()
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    return True
