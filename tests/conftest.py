import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libthindb_gpu.so)")
    config.addinivalue_line("markers", "slow: long-running (full-size workloads)")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def gen_lib():
    from thindb_amd import gen
    gen.lib()
    return gen


@pytest.fixture(scope="session")
def session():
    """One GPU session for every GPU test of the run."""
    from thindb_amd.exec import GpuSession
    s = GpuSession(0)
    yield s
    s.close()
