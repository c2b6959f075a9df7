"""Shared fixtures. Tests marked `gpu` need an MI355X; everything else runs on the CPU."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (MI355X)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import Oracle  # oracle/oracle.py
    return Oracle()


@pytest.fixture(scope="session")
def reflib():
    from oracle import RefLib
    if not RefLib.available():
        pytest.skip("oracle/_ref/libnova_ref.so is not built (needs the reference tree)")
    return RefLib()


@pytest.fixture(scope="session")
def lib():
    from nova_amd import _abi
    return _abi.load()


def bits(a):
    """float32 array as its raw bit patterns, for bit-exact comparisons."""
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)
