"""Shared test setup: markers, paths and circuit fixtures."""

from __future__ import annotations

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: takes more than a few seconds")


def circuit_path(name: str) -> str:
    return os.path.join(ROOT, "greenpeas_amd", "data", "circuits", f"{name}.stim")


def read_circuit(name: str) -> str:
    with open(circuit_path(name)) as f:
        return f.read()


def golden_dem_path(name: str, level: int) -> str:
    return os.path.join(GOLDEN, "dems", f"{name}_L{level}.dem")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    O.lib()
    return O
