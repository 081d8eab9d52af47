import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device; calls the HIP library")


@pytest.fixture(scope="session")
def ctx():
    """One spx context on device 0.  Fails (never falls back) if the HIP library is absent."""
    from spectrum_amd import capi

    capi.load()  # raises if libspectrum_mi355x.so was not built
    c = capi.Context(0)
    print(f"[gpu] {c.info()}", flush=True)
    yield c
    c.close()
