import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


def load_stan_csv(path):
    """Parse a Stan CSV: returns (column names, draws array, comment lines)."""
    names, rows, comments = None, [], []
    with open(path) as f:
        for line in f:
            if line.startswith("#"):
                comments.append(line.rstrip("\n"))
            elif names is None:
                names = line.strip().split(",")
            elif line.strip():
                rows.append([float(v) for v in line.split(",")])
    import numpy as np
    return names, np.array(rows), comments


@pytest.fixture(scope="session")
def golden():
    return GOLDEN
