import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")
    config.addinivalue_line("markers", "slow: large-size parity cases")
    # Build only when nothing is built (the round-end GPU run uses the prebuilt tree).
    from platoanalyze_amd.build import LIB, ORACLE_LIB, build_native, build_oracle

    if not LIB.exists():
        build_native()
    if not ORACLE_LIB.exists():
        build_oracle()


def pytest_collection_modifyitems(config, items):
    # keep CPU runs lean: the oracle uses at most 8 threads here
    os.environ.setdefault("OMP_NUM_THREADS", "8")


@pytest.fixture(scope="session")
def goldens():
    import json

    return json.loads((ROOT / "tests" / "golden" / "reference_goldens.json").read_text())


@pytest.fixture(scope="session")
def box222():
    from platoanalyze_amd.mesh import build_box

    return build_box(2, 2, 2, 1.0, 1.0, 1.0)
