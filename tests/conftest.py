import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and libqem_hip.so")
    config.addinivalue_line("markers", "slow: takes more than a few seconds")


def _announce(item):
    # name the case before any GPU work starts, so a crash is attributable
    if item.get_closest_marker("gpu") is not None:
        sys.stdout.write(f"\n[gpu-case] {item.nodeid}\n")
        sys.stdout.flush()


def pytest_runtest_setup(item):
    _announce(item)
