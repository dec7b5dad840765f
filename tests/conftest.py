import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built librrh.so")


@pytest.fixture(scope="session")
def ctx():
    from research_amd import core

    c = core.Context(0)
    yield c  # released with the process: scenes and images created by tests still use it
