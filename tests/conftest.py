import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, 'tests', 'golden')


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (runs the HIP kernels)')


def load_dat(name):
    """Column data file (header line + rows) -> (header, array)."""
    import numpy as np
    with open(os.path.join(GOLDEN, name)) as f:
        hdr = f.readline().split()
        rows = [l.split() for l in f if l.strip()]
    return hdr, np.array(rows, dtype=float)


@pytest.fixture(scope='session')
def oracle():
    from oracle import pyoracle
    pyoracle.lib()
    return pyoracle
