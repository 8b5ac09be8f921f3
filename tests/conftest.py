import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, 'tvb-hpc_amd'), os.path.join(ROOT, 'oracle'),
          os.path.join(ROOT, 'tests')):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, 'tests', 'golden')


def pytest_configure(config):
    config.addinivalue_line(
        'markers', 'gpu: needs an MI355X device (select with -m gpu)')
    config.addinivalue_line('markers', 'slow: longer CPU test')


@pytest.fixture(scope='session')
def ctx():
    from tvb_hpc import hip
    if hip.device_count() < 1:
        pytest.fail('GPU test collected without a visible GPU')
    return hip.Context.default()
