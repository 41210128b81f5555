import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
PKG = os.path.join(ROOT, 'ddalphaamg-sm_amd')
if PKG not in sys.path:
    sys.path.insert(0, PKG)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (HIP device)')
    config.addinivalue_line('markers', 'slow: longer CPU test')


GOLDEN = os.path.join(ROOT, 'tests', 'golden')


@pytest.fixture(scope='session')
def golden():
    import json

    import numpy as np

    def load(name):
        with open(os.path.join(GOLDEN, name + '.json')) as f:
            meta = json.load(f)
        arr = dict(np.load(os.path.join(GOLDEN, name + '.npz')))
        return meta, arr
    return load
