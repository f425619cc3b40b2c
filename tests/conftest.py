import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: longer CPU test")


def pytest_collection_modifyitems(config, items):
    # GPU tests stay selectable with -m gpu; nothing is skipped silently here.
    pass


@pytest.fixture(scope="session")
def cooling_table():
    from galacticus_amd.cooling import synthetic_cie_table
    return synthetic_cie_table()
