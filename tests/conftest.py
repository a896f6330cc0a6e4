"""Shared fixtures of the test-suite.

CPU tests (`-m "not gpu"`) pin the oracle against golden vectors and closed
forms and check the host code and the C ABI; GPU tests (`-m gpu`) compare
the engine's HIP path with the oracle through the C ABI.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
GOLDEN = Path(__file__).resolve().parent / "golden"
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: statistical or long-running test")


@pytest.fixture(scope="session")
def golden() -> Path:
    return GOLDEN


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle

    oracle.lib()
    return oracle
