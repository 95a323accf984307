"""Test configuration: puts the repo root on sys.path, registers the `gpu`
marker and builds the libraries if they are missing."""
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (runs through libcudasph.so)")


@pytest.fixture(scope="session")
def orc():
    import oracle
    if not oracle.LIB_PATH.exists():
        oracle.build()
    return oracle


@pytest.fixture(scope="session")
def cs():
    import cudasph_amd
    if not cudasph_amd.LIB_PATH.exists():
        from cudasph_amd import build
        build.build()
    return cudasph_amd
