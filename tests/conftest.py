"""Shared pytest configuration.

Tests marked ``gpu`` need a HIP device (MI355X, gfx950); everything else runs
on the CPU.  The CPU oracle under ``oracle/`` is the checker for every parity
test.  With PDQ_PARITY_REPORT=<path> the accuracy numbers of every parity
comparison made in the session are written to <path> as JSON.
"""

import json
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
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def pytest_sessionfinish(session, exitstatus):
    path = os.environ.get("PDQ_PARITY_REPORT")
    mod = sys.modules.get("test_gpu_parity")
    if path and mod is not None and mod.REPORT:
        with open(path, "w") as f:
            json.dump(mod.REPORT, f, indent=0)


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    O.build()
    return O
