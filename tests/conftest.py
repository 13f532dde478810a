import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle():
    from oracle.binding import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def reference():
    from oracle.binding import REF_SO, Reference
    if not os.path.exists(REF_SO):
        pytest.skip("reference host build (oracle/_ref) not present")
    return Reference()


@pytest.fixture(scope="session")
def native():
    from h261enc_cuda_amd import _native
    return _native.load()


def announce(msg: str) -> None:
    """Names the case on stdout before a GPU step (visible if the step dies)."""
    print(f"[case] {msg}", flush=True)
