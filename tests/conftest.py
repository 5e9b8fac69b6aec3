"""Shared test setup: import paths, the `gpu` marker, GPU gating."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "photons-on-mt3000_amd")
ORACLE = os.path.join(ROOT, "oracle")
for p in (PKG, ORACLE, ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a visible MI355X (HIP) device")


def _gpu_visible() -> bool:
    try:
        import photons_p2p
        return photons_p2p.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu_ctx():
    """One library context for the GPU tests; the tests fail (not skip) when
    the library or the device is missing on a GPU run."""
    import photons_p2p
    ctx = photons_p2p.Context(0)
    yield ctx
    ctx.close()
