"""Test configuration: markers and import paths.

GPU tests (``-m gpu``) call the HIP library through its C ABI and compare with
the oracle under ``oracle/`` on the same seeded inputs.
"""
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT, ROOT / "pr-calc_amd", ROOT / "oracle"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X and the built libprc_d2z.so")
    config.addinivalue_line("markers", "slow: larger grids (minutes)")


def announce(msg: str) -> None:
    """Print the case before a GPU step so a failing run names it."""
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()
