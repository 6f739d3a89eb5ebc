"""Shared pytest configuration.

Tests marked ``gpu`` need an MI355X (gfx950) and the in-tree native
libraries; everything else runs on a CPU-only host.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU and the built native libraries")


@pytest.fixture(scope="session")
def oracle():
    from oracle import binding
    binding.lib()
    return binding
