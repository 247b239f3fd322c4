import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, 'tests', 'golden'))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (runs the HIP engine)")
    config.addinivalue_line("markers", "slow: larger parity cases")


@pytest.fixture(scope='session')
def oracle_lib():
    """The CPU oracle (built from oracle/ on demand)."""
    path = os.path.join(ROOT, 'oracle', 'libtc_oracle.so')
    if not os.path.exists(path):
        subprocess.check_call(['make', '-s', '-f', os.path.join(ROOT, 'oracle', 'Makefile')],
                              cwd=ROOT)
    from oracle import oracle
    return oracle


@pytest.fixture(scope='session')
def gpu():
    """The treecorr_amd package with a usable GPU (fails loudly otherwise)."""
    import treecorr_amd
    from treecorr_amd import _lib
    n = _lib.device_count()
    assert n > 0, "GPU test ran without a visible GPU"
    return treecorr_amd
