import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; runs through the C ABI")
    config.addinivalue_line("markers", "slow: long CPU test")


def _gpu_count() -> int:
    try:
        import shapeextraction_amd as sx

        return sx.device_count()
    except Exception:
        return 0


@pytest.fixture(scope="session")
def sx():
    import shapeextraction_amd as sx_mod

    return sx_mod


@pytest.fixture(scope="session")
def gpu(sx):
    """Session fixture for GPU tests: fails loudly when the library cannot see a GPU."""
    n = sx.device_count()
    if n == 0:
        pytest.fail("GPU test selected but no HIP device is visible to libshapes_hip.so")
    print(f"[gpu] {sx.device_name(0)}", flush=True)
    return sx


@pytest.fixture(scope="session")
def O():
    from oracle import oracle

    return oracle


@pytest.fixture(autouse=True)
def _announce(request):
    # Name the case before any GPU work so that a fault can be attributed.
    print(f"\n[case] {request.node.nodeid}", flush=True)
    yield


def rng(seed: int) -> np.random.Generator:
    return np.random.default_rng(seed)
