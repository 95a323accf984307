"""Test configuration: puts the repo root on sys.path, registers the `gpu`
marker and builds the libraries when they are missing or older than their
sources (libcudasph.so records a hash of its sources; the oracle's Makefile
compares times)."""
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
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def cs():
    from cudasph_amd import build
    build.ensure_built()
    import cudasph_amd
    return cudasph_amd
