"""Test configuration: the `gpu` marker and shared fixtures.

CPU tests (`-m "not gpu"`) pin the oracle against closed forms and independent
libraries and check the C ABI exports; GPU tests compare the HIP path with the
oracle on the same inputs.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.build()
    return O


def _have_gpu() -> bool:
    try:
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


@pytest.fixture(scope="session")
def tj():
    """The HIP library; GPU tests fail (not skip) when it cannot be used."""
    import tomoj_amd
    tomoj_amd.load_library()
    if not _have_gpu():
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    return tomoj_amd
