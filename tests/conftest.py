import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "ref: needs the reference built under oracle/_ref")


@pytest.fixture(scope="session")
def oracle_lib():
    from tests import helpers
    return helpers.oracle_lib()


@pytest.fixture(scope="session")
def gpu_lib():
    import turns2d_urans_git_amd as t2d
    return t2d.load_library()
