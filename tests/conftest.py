import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


def _gpu_available():
    try:
        import powspec_amd as P
        return P.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def psh():
    """The HIP library; GPU tests fail (not skip) when it cannot load on a GPU
    box, so a missing extension is loud."""
    import powspec_amd as P
    P.load()
    if P.device_count() < 1:
        pytest.skip("no GPU visible")
    return P


@pytest.fixture(scope="session")
def psh_host():
    """The HIP library for its host-only entry points (no GPU needed)."""
    import powspec_amd as P
    P.load()
    return P
