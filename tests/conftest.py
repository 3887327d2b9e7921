import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, 'tests', 'golden')


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and libssqk.so")


class _Golden:
    def __init__(self):
        self._cache = {}

    def __call__(self, name):
        if name not in self._cache:
            self._cache[name] = dict(np.load(os.path.join(GOLDEN, name + '.npz')))
        return self._cache[name]


@pytest.fixture(scope='session')
def golden():
    return _Golden()


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))
