import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, 'idea.deploy_amd')):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (runs the HIP kernels)')
    config.addinivalue_line('markers', 'slow: long CPU oracle run')


@pytest.fixture(scope='session')
def tenet():
    from idmx.IdpyCode import Tenet
    t = Tenet(0)
    yield t
    t.End()
