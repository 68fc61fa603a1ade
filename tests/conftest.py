import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle_engine():
    from oracle.oracle import OracleEngine
    return OracleEngine(threads=4)


@pytest.fixture(scope="session")
def gpu_engine():
    from minilake_amd.gpu import GpuEngine
    eng = GpuEngine()
    yield eng
    eng.close()
