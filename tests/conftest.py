import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (runs libkpcmi kernels)")
    config.addinivalue_line("markers", "slow: CPU test that takes tens of seconds")


@pytest.fixture(scope="session")
def dcov_fixture():
    with open(os.path.join(GOLDEN, "dcov_gamma_oracle_v1.json")) as fh:
        fx = json.load(fh)
    cases = []
    for c in fx["cases"]:
        c = dict(c)
        c["rx"] = np.array([float.fromhex(v) for v in c["rx"]])
        c["ry"] = np.array([float.fromhex(v) for v in c["ry"]])
        cases.append(c)
    return cases


@pytest.fixture(scope="session")
def dev500():
    return np.load(os.path.join(GOLDEN, "dev500x50_v2.npz"))


def trace_map(level, x, y, S, p):
    return {(int(a), int(b), int(c), str(d)): float(e) for a, b, c, d, e in zip(level, x, y, S, p)}
