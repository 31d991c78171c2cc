"""Shared fixtures: package loader, GPU marker, oracle access."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs a HIP device (MI355X)')


@pytest.fixture(scope='session')
def pkg():
    import __graft_entry__ as ge
    return ge.load_package()


@pytest.fixture(scope='session')
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.fail('gpu-marked test without a visible HIP device')
    return torch.device('cuda:0')
