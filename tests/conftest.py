"""Shared test setup: import paths and the `gpu` marker.

Tests marked ``gpu`` run on an MI355X; everything else runs on the CPU (no
device calls).  The oracle under oracle/ is used only as the checker.
"""
import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
for p in (ROOT / "chymera-uo_amd", ROOT / "oracle", ROOT / "tests" / "golden"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def golden_dir():
    return GOLDEN
