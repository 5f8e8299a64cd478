import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long-running")


def _ensure_built():
    from mfix_amd import LIB_PATH
    if not LIB_PATH.exists():
        from mfix_amd import build
        build.build()


@pytest.fixture(scope="session")
def native_lib():
    _ensure_built()
    from mfix_amd import _lib
    return _lib.load()


@pytest.fixture(scope="session")
def gpu(native_lib):
    """Initialise device 0 once per session; fail loudly when it is absent."""
    from mfix_amd import leq
    leq.init(0)
    yield leq


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.build_oracle()
    return oracle
