import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope='session')
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope='session')
def ref_bresenham():
    from oracle import oracle as O
    mod = O.ref_module('bresenham')
    if mod is None:
        pytest.skip("oracle/_ref/bresenham is not built (reference sources unavailable)")
    return mod


@pytest.fixture(scope='session')
def ref_label():
    from oracle import oracle as O
    mod = O.ref_module('label')
    if mod is None:
        pytest.skip("oracle/_ref/label is not built (reference sources unavailable)")
    return mod


@pytest.fixture(scope='session')
def cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    return torch.device('cuda:0')
