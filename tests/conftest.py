"""Shared test set-up: import paths, the `gpu` marker, case helpers."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "efdcplus8.5_amd"))
sys.path.insert(0, ROOT)

G = 9.81


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (runs the HIP library through the C-ABI)")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle as orc
    orc.build()
    return orc


def wet_index(case):
    return np.arange(2, case.grid.LA + 1) - 1


def eta_err(p_a, p_b):
    """max |delta eta| in metres for two P fields (P = g*eta)."""
    return float(np.max(np.abs(p_a - p_b))) / G
