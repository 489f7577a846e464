import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libmgbfs_hip.so")
    config.addinivalue_line("markers", "slow: long CPU oracle runs, opt-in with MGBFS_SLOW=1")


def pytest_collection_modifyitems(config, items):
    if os.environ.get("MGBFS_SLOW") == "1":
        return
    skip = pytest.mark.skip(reason="slow oracle run; set MGBFS_SLOW=1")
    for item in items:
        if "slow" in item.keywords:
            item.add_marker(skip)
