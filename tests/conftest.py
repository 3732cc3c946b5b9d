import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


@pytest.fixture(scope="session")
def golden():
    import json
    import numpy as np
    d = os.path.join(ROOT, "tests", "golden")
    with open(os.path.join(d, "cases.json")) as f:
        cases = json.load(f)
    arrays = np.load(os.path.join(d, "cases.npz"))
    with open(os.path.join(d, "hashes.json")) as f:
        hashes = json.load(f)
    return cases, arrays, hashes
