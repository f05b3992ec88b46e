import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, 'tests', 'golden')


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (gfx950) GPU and the built HIP library')


def load_golden(name):
    with open(os.path.join(GOLDEN, f'{name}.json')) as fh:
        meta = json.load(fh)
    cells0 = np.load(os.path.join(GOLDEN, f'{name}_cells0.npy'))
    frame = np.load(os.path.join(GOLDEN, f'{name}_last_frame.npy'))
    return meta, cells0, frame


@pytest.fixture(scope='session')
def golden():
    return load_golden
