"""Shared fixtures.  GPU tests are marked ``gpu``; everything else runs on a
CPU-only host (the C ABI library loads there; only its host code is called)."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import __graft_entry__  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


@pytest.fixture(scope="session")
def cr():
    return __graft_entry__.load_package()


@pytest.fixture(scope="session")
def orc():
    return __graft_entry__.load_oracle()
