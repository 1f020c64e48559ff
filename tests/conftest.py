"""Shared test configuration.

Markers: ``gpu`` (needs an MI355X through the HIP library), ``slow``.
"""
import os
import random
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (str(ROOT), str(ROOT / "atlas-q_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP library on cuda:0)")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(autouse=True)
def _seed_everything():
    import torch
    random.seed(0)
    np.random.seed(0)
    torch.manual_seed(0)
    yield


def load_golden(name):
    return dict(np.load(GOLDEN / f"{name}.npz", allow_pickle=False))
