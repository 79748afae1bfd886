"""pytest configuration: the `gpu` marker and shared loaders.

CPU tests (`-m "not gpu"`) cover the oracle against brute force / closed forms
/ the reference's own golden fixtures, the host library, and that the C ABI
libraries load and export what include/*.h declares.  GPU tests (`-m gpu`)
compare the HIP kernels, called through the C ABI, with the oracle.
"""
import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


@pytest.fixture(scope="session")
def g2():
    return importlib.import_module("g2-aimd_amd")


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle


@pytest.fixture(scope="session")
def gpu_exec(g2):
    """One device context for the whole GPU session (fails loudly without a GPU)."""
    ex = g2.Executor(device=0)
    yield ex
    ex.close()
