"""Shared test setup.

Markers: `gpu` tests need a HIP device and libcvortex.so; everything else
runs on the CPU.  The CPU oracle (oracle/liboracle.so) is built here when
missing -- it is the checker, compiled from oracle/*.c with gcc.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X) and the built libcvortex.so")


@pytest.fixture(scope="session")
def orc():
    import oracle
    if not os.path.exists(oracle.LIB_PATH):
        oracle.build()
    return oracle.default()
