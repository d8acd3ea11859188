"""Test configuration: the `gpu` marker, the package loader and shared helpers.

Tests marked `gpu` call the HIP library through its C ABI on a real device; the others run on the
CPU (oracle pins, fixtures, host logic, library loading and symbol export).
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import __graft_entry__ as _ge  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and the built libnlfh.so")


def load_pkg():
    return _ge.load_package()


@pytest.fixture(scope="session")
def pkg():
    return load_pkg()


@pytest.fixture(scope="session")
def gpu_pkg():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    p = load_pkg()
    p.load_library()
    return p
