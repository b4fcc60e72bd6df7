import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")
    config.addinivalue_line("markers", "slow: larger CPU-oracle cases")


def rel_l2(a, b) -> float:
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle

    oracle.load()
    return oracle


@pytest.fixture(scope="session")
def gpu_ctx():
    """An hgx context on cuda:0. The library must be built; there is no CPU fallback."""
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU in this environment")
    from hipgisaxs_amd.api import Context

    ctx = Context(0)
    yield ctx
    ctx.close()
