import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI library)")


@pytest.fixture(scope="session")
def reference_fixtures():
    with open(os.path.join(GOLDEN, "reference_fixtures.json")) as f:
        return json.load(f)


def csr_from_fixture(fx, dtype=np.int32):
    from oracle.oracle import Csr

    arcs = fx["arcs"]
    src = [a[0] for a in arcs]
    dst = [a[1] for a in arcs]
    w = [a[2] for a in arcs]
    return Csr.from_edges(fx["n"], src, dst, w, dtype=dtype)
