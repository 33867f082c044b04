"""Shared fixtures: the oracle (CPU restatement, checker only) and the HIP
library under test."""
import ctypes
import importlib.util
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "_build", "liboracle.so")
PKG_DIR = os.path.join(ROOT, "enblend-enfuse_amd")

if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def load_package():
    """Import enblend-enfuse_amd/ (a directory name that is not an identifier)."""
    if "enblend_enfuse_amd" in sys.modules:
        return sys.modules["enblend_enfuse_amd"]
    spec = importlib.util.spec_from_file_location("enblend_enfuse_amd", os.path.join(PKG_DIR, "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["enblend_enfuse_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def oracle():
    from oracle_binding import load_oracle
    return load_oracle()


@pytest.fixture(scope="session")
def enbl():
    return load_package()


@pytest.fixture(scope="session")
def gpu_ctx(enbl):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    ctx = enbl.Context(0)
    yield ctx
    ctx.close()
