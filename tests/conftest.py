"""Shared test setup: paths, the `gpu` marker, oracle access."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "fomo.jl_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")
    config.addinivalue_line("markers", "slow: long CPU oracle runs (seconds to a minute)")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib("f32")
    oracle.lib("f64")
    return oracle


def _gpu_available() -> bool:
    try:
        import fomo._lib as L
        return L.lib().fomo_device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Skip-free guard: a `gpu` test on a machine without a GPU fails loudly."""
    import fomo._lib as L
    n = L.lib().fomo_device_count()
    assert n > 0, "no HIP device visible; gpu tests must run on an MI355X"
    return n
