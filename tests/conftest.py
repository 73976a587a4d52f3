import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X) and liblfr.so")


@pytest.fixture(scope="session")
def ctx():
    from lfric_core_amd.api import Context
    c = Context(0)
    yield c
    c.close()
