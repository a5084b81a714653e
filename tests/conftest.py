"""Shared test setup: the `gpu` marker, and loaders for the oracle (CPU checker) and the
product library (libc3d_hip.so)."""
import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


def _load_module(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def product():
    """c3d-2015_amd/c3d_hip.py (the ctypes front end of libc3d_hip.so)."""
    if "c3d_hip" in sys.modules:
        return sys.modules["c3d_hip"]
    return _load_module("c3d_hip", os.path.join(ROOT, "c3d-2015_amd", "c3d_hip.py"))


@pytest.fixture(scope="session")
def P():
    return product()


@pytest.fixture(scope="session")
def O():
    import oracle
    return oracle


@pytest.fixture(scope="session")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")
