"""Shared test setup.

Markers: `gpu` tests need an MI355X (they run the HIP library); everything
else runs on the CPU: the oracle against golden vectors and closed forms, the
host-side initializer, and the C-ABI library's symbol table.
"""
import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (runs HIP kernels)")


@pytest.fixture(scope="session")
def picmod():
    return importlib.import_module("0buf_amd")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.load()
    return oracle
