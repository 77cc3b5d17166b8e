import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "homotopy-continuation-tracker-on-gpu_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: longer CPU test")


def gpu_available():
    try:
        import hcx
        return hcx.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def hcx_mod():
    import hcx
    hcx.lib()
    return hcx


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def gpu(hcx_mod):
    """Fails (not skips) when marked GPU tests run without a device: the
    product path must not fall back to anything else."""
    assert hcx_mod.device_count() > 0, "no HIP device visible to libhcx"
    import torch
    assert torch.cuda.is_available()
    return True
