"""Shared fixtures.  Tests marked `gpu` need an MI355X and the built
libmtrt.so; everything else runs on the CPU (oracle, host layer, ABI)."""
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct MI355X (gfx950) and libmtrt.so")


@pytest.fixture(scope="session")
def tracer():
    """One device context for the whole GPU session (fails loudly without a GPU)."""
    from master_thesis_amd.tracer import Tracer
    t = Tracer(0)
    yield t
    t.close()
