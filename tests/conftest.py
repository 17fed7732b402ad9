import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "ref: needs the reference harness oracle/_ref (built "
                                       "from the reference checkout)")


def pytest_collection_modifyitems(config, items):
    import oracles
    if not oracles.ref_available():
        skip = pytest.mark.skip(reason="oracle/_ref not built (no reference checkout)")
        for it in items:
            if "ref" in it.keywords:
                it.add_marker(skip)


@pytest.fixture(scope="session")
def orc():
    import oracles
    return oracles.oracle()


@pytest.fixture(scope="session")
def refl():
    import oracles
    return oracles.ref()


@pytest.fixture(scope="session")
def ctx():
    """One GPU context for the whole GPU test session."""
    from cookiedough_amd.device import Context
    c = Context(int(os.environ.get("CKD_DEVICE", "0")))
    yield c
    c.close()
