import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libsis_gpu.so")
    config.addinivalue_line("markers", "ref: needs the reference build in oracle/_ref")
    config.addinivalue_line("markers", "slow: takes more than ~20 s")


def pytest_collection_modifyitems(config, items):
    from oracle import refdrv
    if not refdrv.available():
        skip = pytest.mark.skip(reason="oracle/_ref (reference Fortran build) not available")
        for it in items:
            if "ref" in it.keywords:
                it.add_marker(skip)
