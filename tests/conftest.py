"""Shared test setup: the `gpu` marker, import paths and common helpers."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

SUITE = ["path1000", "star100k", "complete200", "two_components", "empty1000", "grid100x100", "rmat12",
         "rmat12_directed"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; run with -m gpu")


def sample_sources(n):
    """First, middle, last vertex plus three LCG-random ones (tests/test_main.cpp:286-291)."""
    s = [0, n // 2, n - 1]
    x = 12345
    for _ in range(3):
        x = (x * 6364136223846793005 + 1442695040888963407) & (2**64 - 1)
        s.append((x >> 33) % n)
    return s


def hub(g):
    d = g.degrees()
    return int(np.argmax(d)) if len(d) else 0


@pytest.fixture(scope="session")
def suite_graphs():
    import graph_analytics_applesilicon_amd as G
    return {name: G.Graph.named(name) for name in SUITE}
