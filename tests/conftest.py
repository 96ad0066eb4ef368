import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GRAPHS = os.path.join(ROOT, "tests", "golden", "graphs")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle_py
    if not os.path.exists(oracle_py.LIB_PATH):
        oracle_py.build()
    oracle_py.lib()
    return oracle_py


def graph_path(name: str) -> str:
    return os.path.join(GRAPHS, name)
