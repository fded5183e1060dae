import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
for p in (REPO, REPO / "emory-cci-fast-ia-gsoc_amd", REPO / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: full-size workloads")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def nh():
    import nscale_hip
    nscale_hip._lib.load()
    return nscale_hip


@pytest.fixture(scope="session")
def gpu(nh):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return nh.default_context(0)
